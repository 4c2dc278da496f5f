#!/usr/bin/env bash
# Builds the reference's force path from its own C sources into oracle/_ref/.
#
# TEST INFRASTRUCTURE ONLY.  Outputs (binaries) go to oracle/_ref/ only,
# which git ignores.  The sources are read from the reference checkout
# (MSPMS_REF_DIR, default /root/reference) and compiled through a scratch
# copy under /tmp that receives two edits, then deleted:
#   1. src/frclong.c:86, the j-molecule atom offset:
#        jabs = jsum_atom + (jStartMole==DYNAMIC_ID ? jAtom : 0);
#      becomes
#        jabs = (jStartMole==DYNAMIC_ID ? iabs + iAtom : jsum_atom);
#      (unpatched, every species with >2 molecules gets r=0 self pairs and
#      NaN forces).
#   2. src/molevars.h:6-13 array limits raised to 131072 atoms / molecules
#      and 262144 bonded terms so 12,288- and 100,000-site systems fit.
# Nothing is stubbed: dftmcff.c (needs GSL) and mspms2.c (main) are left
# out, and the references into the absent Fortran grid library, GSL and the
# metal-cluster force field stay unresolved at link time.  ref_harness.c
# refuses the inputs that would reach them.
set -euo pipefail
REF=${MSPMS_REF_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/frclong.c" ]; then
    echo "build_ref.sh: no reference sources at $REF; nothing built" >&2
    exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/mspms_ref_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
cp "$REF"/src/*.c "$REF"/src/*.h "$TMP"/
rm -f "$TMP/dftmcff.c" "$TMP/mspms2.c"

OLD='jabs = jsum_atom + (jStartMole==DYNAMIC_ID ? jAtom : 0);'
NEW='jabs = (jStartMole==DYNAMIC_ID ? iabs + iAtom : jsum_atom);'
grep -qF "$OLD" "$TMP/frclong.c"
sed -i "s/jabs = jsum_atom + (jStartMole==DYNAMIC_ID ? jAtom : 0);/$NEW/" "$TMP/frclong.c"
grep -qF "$NEW" "$TMP/frclong.c"

sed -i -e 's/^#define NATOM_MAX.*/#define NATOM_MAX\t131072/' \
       -e 's/^#define NMOLE_MAX.*/#define NMOLE_MAX\t131072/' \
       -e 's/^#define NBOND_MAX.*/#define NBOND_MAX\t262144/' \
       -e 's/^#define NANGLE_MAX.*/#define NANGLE_MAX\t262144/' \
       -e 's/^#define NDIH_MAX.*/#define NDIH_MAX\t262144/' \
       -e 's/^#define NIMP_MAX.*/#define NIMP_MAX\t262144/' \
       -e 's/^#define NNBP_MAX.*/#define NNBP_MAX\t262144/' "$TMP/molevars.h"
grep -q 'NATOM_MAX.131072' "$TMP/molevars.h"

CFLAGS="-std=c99 -O2 -fcommon -w"
(
    cd "$TMP"
    for f in *.c; do gcc $CFLAGS -c "$f" -o "${f%.c}.o"; done
    gcc $CFLAGS -I"$TMP" -c "$HERE/ref_harness.c" -o harness.o
    gcc -fcommon -o "$OUT/mspms_ref_harness" harness.o $(ls *.o | grep -v '^harness.o$') \
        -lm -Wl,--unresolved-symbols=ignore-all
    # objects for the drop-in integration harness (reference callers + the
    # engine's frclong/frcshort); frclong.o / frcshort.o are left out.
    ar rcs "$OUT/libmspms_ref_callers.a" $(ls *.o | grep -v -e '^harness.o$' -e '^frclong.o$' -e '^frcshort.o$')
    cp harness.o "$OUT/ref_harness.o"
    # Term harness: the reference's own frclong / frcshort and callers with
    # the engine's term-level functions (mspms2_amd/csrc/dropin_terms.c over
    # libmspms_gpu.so's host term code) in place of ljfrc.c, elecfrc.c and
    # the bonded-term files; sffrc.c stays for its start-up readers, and its
    # cal_sf_* / get_values_from_grid yield to the engine's (linked first).
    PKG="$HERE/../mspms2_amd"
    if [ -f "$PKG/_build/dropin_terms.o" ] && [ -f "$PKG/libmspms_gpu.so" ]; then
        gcc -fcommon -o "$OUT/mspms_terms_harness" harness.o "$PKG/_build/dropin_terms.o" \
            "$PKG/_build/dropin_globals.o" \
            $(ls *.o | grep -v -e '^harness.o$' -e '^ljfrc.o$' -e '^elecfrc.o$' \
                  -e '^bndfrc.o$' -e '^aglfrc.o$' -e '^dihfrc.o$' -e '^impfrc.o$') \
            -L"$PKG" -lmspms_gpu -lm -Wl,--allow-multiple-definition \
            -Wl,-rpath,'$ORIGIN/../../mspms2_amd' -Wl,-rpath-link,/opt/rocm/lib \
            -Wl,--unresolved-symbols=ignore-in-object-files
        echo "build_ref.sh: built $OUT/mspms_terms_harness"
    fi
    # layout probe of SAMPLE_MOLECULE for the drop-in ABI check
    cat > layout.c <<'EOF'
#include <stdio.h>
#include <stddef.h>
#include "mspms2.h"
int main(void){
  printf("sizeof %zu\n", sizeof(SAMPLE_MOLECULE));
  printf("type %zu\nghost_type %zu\ninterp_type %zu\naw %zu\nepsilon %zu\nsigma %zu\ncharge %zu\n",
    offsetof(SAMPLE_MOLECULE,type), offsetof(SAMPLE_MOLECULE,ghost_type), offsetof(SAMPLE_MOLECULE,interp_type),
    offsetof(SAMPLE_MOLECULE,aw), offsetof(SAMPLE_MOLECULE,epsilon), offsetof(SAMPLE_MOLECULE,sigma), offsetof(SAMPLE_MOLECULE,charge));
  printf("mw %zu\nnatom %zu\nnbond %zu\nnangle %zu\nndih %zu\nnimp %zu\nnnbp %zu\n",
    offsetof(SAMPLE_MOLECULE,mw), offsetof(SAMPLE_MOLECULE,natom), offsetof(SAMPLE_MOLECULE,nbond),
    offsetof(SAMPLE_MOLECULE,nangle), offsetof(SAMPLE_MOLECULE,ndih), offsetof(SAMPLE_MOLECULE,nimp), offsetof(SAMPLE_MOLECULE,nnbp));
  printf("bnd_idx %zu\nbnd_type %zu\nbnd_para_0 %zu\nbnd_para_1 %zu\nbnd_para_2 %zu\n",
    offsetof(SAMPLE_MOLECULE,bnd_idx), offsetof(SAMPLE_MOLECULE,bnd_type), offsetof(SAMPLE_MOLECULE,bnd_para_0),
    offsetof(SAMPLE_MOLECULE,bnd_para_1), offsetof(SAMPLE_MOLECULE,bnd_para_2));
  printf("agl_idx %zu\nagl_type %zu\nagl_para_1 %zu\nagl_para_5 %zu\nisAngle_unique %zu\n",
    offsetof(SAMPLE_MOLECULE,agl_idx), offsetof(SAMPLE_MOLECULE,agl_type), offsetof(SAMPLE_MOLECULE,agl_para_1),
    offsetof(SAMPLE_MOLECULE,agl_para_5), offsetof(SAMPLE_MOLECULE,isAngle_unique));
  printf("dih_idx %zu\ndih_type %zu\ndih_para_1 %zu\ndih_para_4 %zu\nisDih_unique %zu\n",
    offsetof(SAMPLE_MOLECULE,dih_idx), offsetof(SAMPLE_MOLECULE,dih_type), offsetof(SAMPLE_MOLECULE,dih_para_1),
    offsetof(SAMPLE_MOLECULE,dih_para_4), offsetof(SAMPLE_MOLECULE,isDih_unique));
  printf("imp_idx %zu\nimp_type %zu\nimp_para_0 %zu\nimp_para_1 %zu\nnbp_idx %zu\n",
    offsetof(SAMPLE_MOLECULE,imp_idx), offsetof(SAMPLE_MOLECULE,imp_type), offsetof(SAMPLE_MOLECULE,imp_para_0),
    offsetof(SAMPLE_MOLECULE,imp_para_1), offsetof(SAMPLE_MOLECULE,nbp_idx));
  return 0;
}
EOF
    gcc $CFLAGS -I"$TMP" layout.c -o layout
    ./layout > "$OUT/sample_molecule_layout.txt"
)
echo "build_ref.sh: built $OUT/mspms_ref_harness"

# Integration harness: the reference's start-up path and callers with the
# engine's drop-in frclong/frcshort (libmspms_dropin.a + libmspms_gpu.so)
# linked in place of src/frclong.c and src/frcshort.c.
PKG="$HERE/../mspms2_amd"
if [ -f "$PKG/libmspms_dropin.a" ] && [ -f "$PKG/libmspms_gpu.so" ]; then
    gcc -fcommon -o "$OUT/mspms_dropin_harness" "$OUT/ref_harness.o" "$PKG/libmspms_dropin.a" \
        "$OUT/libmspms_ref_callers.a" -L"$PKG" -lmspms_gpu -lm \
        -Wl,-rpath,'$ORIGIN/../../mspms2_amd' -Wl,-rpath-link,/opt/rocm/lib \
        -Wl,--unresolved-symbols=ignore-in-object-files
    echo "build_ref.sh: built $OUT/mspms_dropin_harness"
fi

# The reference's offline solid-fluid grid generator (tools/sfgridgen.c,
# plain C; it reads its input file and writes out.sfgridgen in the current
# directory), compiled as it stands.
if [ -f "$REF/tools/sfgridgen.c" ]; then
    gcc -O2 -w -o "$OUT/sfgridgen" "$REF/tools/sfgridgen.c" -lm
    echo "build_ref.sh: built $OUT/sfgridgen"
fi
