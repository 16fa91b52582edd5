#!/bin/bash
# Rebuilds tests/golden/kkt_20260904.json from kkt_instances.cpp (host g++ only).
set -e
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
g++ -O2 -std=c++17 "$here/kkt_instances.cpp" -o "$tmp/kkt_instances"
"$tmp/kkt_instances" > "$here/kkt_20260904.json"
rm -r "$tmp"
