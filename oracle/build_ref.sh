#!/bin/bash
# Builds the reference router's CPU program (pcbroute, from its own sources under /root/reference/src; no GPU back end,
# no ONNX Runtime) into oracle/_ref/pcbroute. The tests use its --eval (the reference's judge of a session file
# against a DSN: open connections and clearance violations) on sessions this repository's router writes. A check
# only: nothing in the package runs it. Skips when the reference sources are absent.
set -e
SRC=${PCBROUTE_REF:-/root/reference}/src
OUT=$(cd "$(dirname "$0")" && pwd)/_ref
[ -d "$SRC" ] || { echo "no reference sources at $SRC"; exit 0; }
mkdir -p "$OUT/obj"
objs=()
for f in "$SRC"/*.cpp; do
  o="$OUT/obj/$(basename "${f%.cpp}").o"
  g++ -O2 -std=c++17 -w -c -o "$o" "$f" &
  objs+=("$o")
done
wait
g++ -O2 -o "$OUT/pcbroute.tmp" "${objs[@]}" -pthread
mv "$OUT/pcbroute.tmp" "$OUT/pcbroute"
rm -rf "$OUT/obj"
echo "built $OUT/pcbroute"
