#!/usr/bin/env bash
# Writes the reference-CPU golden fixtures under tests/golden/.
#
# Runs in the build container only (it needs the reference sources at
# $REF, default /root/reference): builds oracle/_ref/ref_harness from those
# sources in place (oracle/Makefile) and lets it call the reference functions
# on small cases. The fixtures are data: inputs and the reference's outputs.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
make -s -C "$ROOT/oracle" ref
H="$ROOT/oracle/_ref/ref_harness"
"$H" case3d "$HERE/cavity3d_n6_v5" 6 5
"$H" case3d "$HERE/cavity3d_n9_v3" 9 3
"$H" case2d "$HERE/cavity2d_n10_v8" 10 8
"$H" cube   "$HERE/cube_n12_v5" 12 5 1.0
# restart files written by the reference's ckpt::save, with the fields they hold
"$H" ckpt "$HERE/ckpt_cavity3d_n6_v5" cavity3d 6 5
"$H" ckpt "$HERE/ckpt_cube_n8_v4" cube 8 4
# three CPU steps with the MOOD limiter from opposing streams (u_x = +-300 m/s)
"$H" mood "$HERE/mood_cavity3d_n7_v6" cavity3d 7 6 2 3 300 0
"$H" mood "$HERE/mood_cavity2d_n12_v8" cavity2d 12 8 2 3 300 0
# 2-D transport of opposing streams (u_x = +-300 m/s): the quadratic
# reconstruction goes negative where they meet and the positivity repair fires
"$H" pos2d "$HERE/pos_cavity2d_n12_v8" 12 8 300 270 0.5
# cloud management after a body move (remove, merge, refill; 3-D bodies
# are removed and merged, the 2-D ones also refilled)
"$H" cloud "$HERE/cloud_sphere_n12_v3" sphere 12 3 2.0 1.3 0.3
"$H" cloud "$HERE/cloud_cube_n12_v3" cube 12 3 1.5 1.1 0.3
"$H" cloud "$HERE/cloud_circle_n24_v6" circle 24 6 2.0 1.4 0.3
"$H" cloud "$HERE/cloud_square_n24_v6" square 24 6 2.0 1.2 0.3
# the cloud of each body case as main() builds it (embedding, wall state,
# parabolic lid, initial state) and its body
"$H" bodycase "$HERE/body_cube_n12_v4" cube 12 4 2.0
"$H" bodycase "$HERE/body_sphere_n12_v4" sphere 12 4 2.0
"$H" bodycase "$HERE/body_circle_n24_v6" circle 24 6 2.0
"$H" bodycase "$HERE/body_square_n24_v6" square 24 6 2.0
# CPU steps with a free body in opposing streams (u_x = +-300 m/s): the
# body state after every step and the cloud at the end
"$H" bodysteps "$HERE/bodysteps_cube_n12_v4" cube 12 4 2.0 40 300 1
"$H" bodysteps "$HERE/bodysteps_circle_n24_v6" circle 24 6 2.0 40 300 1
# the reference's VTK and CSV writers on a perturbed 2-D and 3-D cloud
"$H" writers "$HERE/writers_cavity3d_n6_v4" cavity3d 6 4
"$H" writers "$HERE/writers_cavity2d_n8_v4" cavity2d 8 4
rm -f "$HERE"/ckpt_*/neigh_flat.npy "$HERE"/ckpt_*/nneigh.npy
( cd "$HERE" && find . -name '*.npy' -o -name 'meta.json' -o -name '*.ckpt' | sort | xargs sha256sum > MANIFEST.sha256 )
echo "fixtures written under $HERE"
