#!/usr/bin/env bash
# Build the reference CRM phase routines, compiled unmodified from the reference tree, plus
# the harness in this directory, into oracle/_ref/libcrmref_<class>.so.
#
# Usage: build_ref.sh NX NY NZ DX DT [NX_RAD NY_RAD]
# Env:   CRM_REF_DIR (default /root/reference), FLANG (default ROCm flang), OPT (default -O2),
#        TAG (appended to the library name, for builds with other OPT flags)
#
# Test infrastructure only: the library is the oracle's reference leg and the CPU baseline.
# Nothing from the reference tree is copied; objects go to oracle/_ref/<class>/ only.
set -euo pipefail
NX=$1; NY=$2; NZ=$3; DX=$4; DT=$5
NXR=${6:-$NX}; NYR=${7:-$NY}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
R=${CRM_REF_DIR:-/root/reference}
FL=${FLANG:-/opt/rocm/lib/llvm/bin/flang}
OPT=${OPT:--O2}
TAG=${TAG:-}
if [ ! -d "$R/components/cam/src/physics/crm" ]; then
  echo "reference tree not found at $R" >&2; exit 3
fi
if [ "$NY" -gt 1 ]; then Y3=1; else Y3=0; fi
CLASS=${NX}x${NY}x${NZ}_dx${DX}_dt${DT}
if [ "$NXR" != "$NX" ] || [ "$NYR" != "$NY" ]; then CLASS=${CLASS}_r${NXR}x${NYR}; fi
B=$OUT/$CLASS$TAG
mkdir -p "$B"
C=$R/components/cam/src/physics/crm
CAM=$R/components/cam/src/physics/cam
SH=$R/cime/src/share/util
DEFS="-cpp -DCRM -Dsam1mom -D_MPDATA -DYES3DVAL=$Y3 -DCRM_NX=$NX -DCRM_NY=$NY -DCRM_NZ=$NZ \
 -DCRM_NX_RAD=$NXR -DCRM_NY_RAD=$NYR -DCRM_DX=$DX -DCRM_DT=$DT"
# Dependency order of the modules on the path (crm_module.F90 itself is not built; see the
# harness header for why).
FILES="$SH/shr_kind_mod.F90 $SH/shr_const_mod.F90 $C/openacc_utils.F90 $C/params.F90
 $CAM/crmdims.F90 $C/domain.F90 $C/ADV_MPDATA/advection.F90 $C/grid.F90 $C/vars.F90
 $C/task_util_NOMPI.F90 $C/utils.F90 $C/sat.F90 $C/MICRO_SAM1MOM/micro_params.F90
 $C/MICRO_SAM1MOM/cloud.F90 $C/MICRO_SAM1MOM/precip_init.F90 $C/MICRO_SAM1MOM/precip_proc.F90
 $C/MICRO_SAM1MOM/microphysics.F90 $C/SGS_TKE/shear_prod2D.F90 $C/SGS_TKE/shear_prod3D.F90
 $C/SGS_TKE/tke_full.F90 $C/SGS_TKE/diffuse_mom2D.F90 $C/SGS_TKE/diffuse_mom3D.F90
 $C/SGS_TKE/diffuse_mom.F90 $C/SGS_TKE/diffuse_scalar2D.F90 $C/SGS_TKE/diffuse_scalar3D.F90
 $C/SGS_TKE/diffuse_scalar.F90 $C/crmtracers.F90 $C/fftpack5_1d.F90 $C/scalar_momentum.F90
 $C/SGS_TKE/sgs.F90 $C/bound_exchange.F90 $C/periodic.F90 $C/boundaries.F90 $C/bound_duvdt.F90
 $C/abcoefs.F90 $C/zero.F90 $C/kurant.F90 $C/buoyancy.F90 $C/forcing.F90 $C/damping.F90
 $C/ice_fall.F90 $C/crmsurface.F90 $C/coriolis.F90 $C/advect2_mom_xy.F90 $C/advect2_mom_z.F90
 $C/advect_mom.F90 $C/fft.F90 $C/press_rhs.F90 $C/press_grad.F90 $C/pressure.F90 $C/adams.F90
 $C/uvw.F90 $C/ADV_MPDATA/advect_scalar2D.F90 $C/ADV_MPDATA/advect_scalar3D.F90
 $C/ADV_MPDATA/advect_scalar.F90 $C/advect_all_scalars.F90 $C/diagnose.F90 $CAM/RNG_MT.F90
 $C/random.F90 $C/setperturb.F90 $C/setparm.F90 $C/task_init.F90"
OBJS=""
cd "$B"
for f in $FILES; do
  o=$(basename "$f" .F90).o
  "$FL" $DEFS -w $OPT -fPIC -c "$f" -o "$o"
  OBJS="$OBJS $o"
done
"$FL" $DEFS -w $OPT -fPIC -c "$HERE/crm_ref_harness.F90" -o crm_ref_harness.o
gcc -O2 -fPIC -c "$C/gammafff.c" -o gammafff.o
"$FL" -shared -o "$OUT/libcrmref_${CLASS}${TAG}.so" $OBJS crm_ref_harness.o gammafff.o
echo "$OUT/libcrmref_${CLASS}${TAG}.so"
