#!/bin/bash
# Build the reference's own CPU LJ+LK op (tmol_ljlk::ljlk_pose_scores and
# ljlk_rotamer_scores), its elec op (tmol_elec::elec_pose_scores), its
# lk_ball ops (tmol_lk_ball::gen_pose_waters, lk_ball_pose_score) and its
# all-pairs-shortest-paths op (tmol_apsp::apsp_op)
# from the sources where they lie under the reference checkout, into
# oracle/_ref/.
# Test infrastructure only: used to cross-check the C oracle and as the CPU
# baseline in bench.py.  Needs only g++, the installed torch and the headers
# the reference vendors (tmol/extern: Eigen, moderngpu).  Nothing is copied.
set -euo pipefail
R=${TMOL_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref/tmol_ljlk_ref.so"
if [ ! -d "$R/tmol/score/ljlk/potentials" ]; then
  echo "reference checkout not found at $R; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
TI=$(python3 -c "import torch.utils.cpp_extension as c; print(' '.join('-I'+p for p in c.include_paths()))")
TL=$(python3 -c "import torch, os; print(os.path.join(os.path.dirname(torch.__file__), 'lib'))")
ABI=$(python3 -c "import torch; print(int(torch._C._GLIBCXX_USE_CXX11_ABI))")
PYI=$(python3 -c "import sysconfig; print(sysconfig.get_paths()['include'])")
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_ljlk_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/score/ljlk/potentials/compiled.ops.cpp" \
  "$R/tmol/score/ljlk/potentials/ljlk_pose_score.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# the reference's all-pairs-shortest-paths op (tmol_apsp::apsp_op), which
# turns the connection-point graph into inter_block_bondsep
OUT2="$HERE/_ref/tmol_apsp_ref.so"
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_apsp_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/pose/compiled/apsp_vestibule.ops.cpp" \
  "$R/tmol/pose/compiled/apsp.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT2.tmp"
mv "$OUT2.tmp" "$OUT2"
echo "built $OUT2"
# the reference's elec op (tmol_elec::elec_pose_scores, elec_rotamer_scores)
OUT3="$HERE/_ref/tmol_elec_ref.so"
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_elec_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/score/elec/potentials/compiled.ops.cpp" \
  "$R/tmol/score/elec/potentials/elec_pose_score.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT3.tmp"
mv "$OUT3.tmp" "$OUT3"
echo "built $OUT3"
# the reference's lk_ball ops (tmol_lk_ball::gen_pose_waters,
# lk_ball_pose_score, lk_ball_rotamer_score)
OUT4="$HERE/_ref/tmol_lk_ball_ref.so"
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_lk_ball_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/score/lk_ball/potentials/compiled.ops.cpp" \
  "$R/tmol/score/lk_ball/potentials/lk_ball_pose_score.cpu.cpp" \
  "$R/tmol/score/lk_ball/potentials/gen_pose_waters.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT4.tmp"
mv "$OUT4.tmp" "$OUT4"
echo "built $OUT4"
# the reference's hbond ops (tmol_hbond::gen_hbond_bases, hbond_pose_scores,
# hbond_rotamer_scores)
OUT5="$HERE/_ref/tmol_hbond_ref.so"
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_hbond_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/score/hbond/potentials/compiled.ops.cpp" \
  "$R/tmol/score/hbond/potentials/hbond_pose_score.cpu.cpp" \
  "$R/tmol/score/hbond/potentials/gen_hbond_bases.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT5.tmp"
mv "$OUT5.tmp" "$OUT5"
echo "built $OUT5"
# the reference's ingest ops (tmol/io/details/compiled: gen_pose_leaf_atoms,
# resolve_his_taut)
OUT6="$HERE/_ref/tmol_io_ref.so"
g++ -O3 -w -std=c++17 -shared -fPIC -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -DTORCH_EXTENSION_NAME=tmol_io_ref \
  -I"$R" -I"$R/tmol/extern" $TI -I"$PYI" \
  "$R/tmol/io/details/compiled/compiled.ops.cpp" \
  "$R/tmol/io/details/compiled/gen_pose_leaf_atoms.cpu.cpp" \
  "$R/tmol/io/details/compiled/resolve_his_taut.cpu.cpp" \
  -L"$TL" -ltorch -ltorch_cpu -lc10 -Wl,-rpath,"$TL" \
  -o "$OUT6.tmp"
mv "$OUT6.tmp" "$OUT6"
echo "built $OUT6"
