#!/bin/bash
# Build the reference Simson `bla` (Fortran) from /root/reference/bla for one
# grid size, into oracle/_ref/.  Test infrastructure only.
#
#   oracle/build_ref.sh NX NY NZ NTHREAD OMP PRESSURE OUTNAME
#
# The reference fixes the grid at compile time in par.f and its README
# instructs users to edit par.f before compiling; this script does that on a
# scratch copy under /tmp (never inside the repository).  Only the binary
# lands in oracle/_ref/OUTNAME.
#
# Compiler: ROCm amdflang with the flags of bla/Makefile's gfortran target
# (-fdefault-real-8 -cpp -O3 [-fopenmp -DOPENMP]).  gfortran is not installed
# here.  amdflang rejects constant out-of-bounds subscripts in argument lists
# of dead scalar>0 code that gfortran accepts, so three work-array declarations
# are widened on the scratch copy (extents only, no arithmetic changes):
#   bla.f     f(...,5+2*scalar,...)    -> max(7,5+2*scalar)
#   bla.f     bis(...,6+2*scalar,...)  -> max(8,6+2*scalar)
#   cbflow.f  bu1/bu2(nxp/2+1,nyp,3+scalar) -> assumed size (*)
# With SMAG=1 in the environment the eddy-viscosity LES models are compiled
# in: bla.f's compile-time flag iiles_smag = iiles*0 becomes iiles*1, the
# switch bla.f:471 documents for allocating the dynamic-procedure arrays.
set -euo pipefail

if [ $# -ne 7 ]; then
    echo "usage: $0 NX NY NZ NTHREAD OMP(yes|no) PRESSURE(0|1) OUTNAME" >&2
    exit 2
fi
NX=$1; NY=$2; NZ=$3; NTHREAD=$4; OMP=$5; PRESSURE=$6; OUT=$7
REF=${SIMSON_REFERENCE:-/root/reference}/bla
HERE=$(cd "$(dirname "$0")" && pwd)
DEST=$HERE/_ref
FC=${FC:-amdflang}

if [ ! -d "$REF" ]; then
    echo "reference sources not found at $REF" >&2
    exit 3
fi
mkdir -p "$DEST"
SCRATCH=$(mktemp -d /tmp/simson_ref.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
cp "$REF"/*.f "$SCRATCH"/

cd "$SCRATCH"
sed -i -E "s/^ +parameter \(nx=[0-9]+,ny=[0-9]+,nz=[0-9]+\)/      parameter (nx=$NX,ny=$NY,nz=$NZ)/" par.f
sed -i -E "s/^ +parameter \(nthread=[0-9]+\)/      parameter (nthread=$NTHREAD)/" par.f
sed -i -E "s/^ +parameter \(pressure=[0-9]+\)/      parameter (pressure=$PRESSURE)/" par.f
grep -q "nx=$NX,ny=$NY,nz=$NZ" par.f
grep -q "nthread=$NTHREAD)" par.f
grep -q "pressure=$PRESSURE)" par.f

sed -i -E 's/^( +real f  \(nx\/2\*mbz\*nyp,)5\+2\*scalar(,nthread\))/\1max(7,5+2*scalar)\2/' bla.f
sed -i -E 's/^( +real bis\(nx\/2\*mbz\*nyp,)6\+2\*scalar(,nthread\))/\1max(8,6+2*scalar)\2/' bla.f
sed -i -E 's/^( +real bu1\(nxp\/2\+1,nyp,)3\+scalar(\),bu2\(nxp\/2\+1,nyp,)3\+scalar\)/\1*\2*)/' cbflow.f
grep -q "max(7,5+2\*scalar)" bla.f
grep -q "max(8,6+2\*scalar)" bla.f
grep -q "bu1(nxp/2+1,nyp,\*),bu2(nxp/2+1,nyp,\*)" cbflow.f
if [ "${SMAG:-0}" = "1" ]; then
    sed -i -E 's/^( +integer,parameter :: iiles_smag = iiles)\*0$/\1*1/' bla.f
    grep -q "iiles_smag = iiles\*1" bla.f
fi

# ppar.f includes 'latest_revision', which bla/Makefile:349-350 generates
# before compiling; run the same generation step here.
echo "      write(*,*) 'Latest Revision in Repository: " \
    "$(grep LastChangedRevision ./*.f | cut -d ' ' -f 3 | sort -n | tail -n 1)" \
    "'" > latest_revision

OBJS="bla ppar preprbl prhs getdt wall_time rkcoeff stopnow rescale part_ini
rparambl comment bflow step dstep xlim blfou blshift cbflow splfre rbflow
bfvort rbla rdiscbl rwavebl lenbox rstreakbl rwavesbl wdiscbl fft2db fft2df
cvecfft_acc boxamp boxcfl wampbl wcflbl wplbl wplbl_serial boxext plnmin
plnmax wextbl boxcflbf wxys boxxys getxy getxz putxy putxz evald evalbc solvbc
cwallbc linearbl nonlinbl advection part_adv fring locf gtrip trip tzran ffun
xysh xzsh fshift init_filter bandsolver hpf_y lpf_y rparamles les sgs_stress
filter_cutoff mhd_poi fringp linearp nonlinp wdiscp rparamwallrough
auxwallrough steadywallrough updatewallrough jetinit bfvort3 pre_nlin_3d
arpck2bla bla2arpck dcheb icheb d2cheb rdcheb setmat trid setmacrhs setmac
tridc crhsc cubip ctim cflush time_string corrd corrp corrd2 ran2 cnxtmp
spline spliet"

FLAGS="-fdefault-real-8 -cpp -O3 -DLINUX_GFORTRAN"
if [ "$OMP" = "yes" ]; then
    FLAGS="$FLAGS -fopenmp -DOPENMP"
fi
JOBS=${REF_BUILD_JOBS:-8}
echo $OBJS | tr ' ' '\n' | xargs -P "$JOBS" -I{} $FC $FLAGS -c {}.f -o {}.o
$FC $FLAGS -o "$DEST/$OUT" $(for o in $OBJS; do echo "$o.o"; done)
echo "built $DEST/$OUT (nx=$NX ny=$NY nz=$NZ nthread=$NTHREAD omp=$OMP pressure=$PRESSURE smag=${SMAG:-0})"
