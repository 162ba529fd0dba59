#!/bin/sh
# Regenerates mrg32k3a_fixtures.json from rocRAND's host engine.
set -e
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
hipcc -O1 -o "$tmp/gen" "$here/gen_mrg32k3a_fixtures.cpp"
"$tmp/gen" > "$here/mrg32k3a_fixtures.json"
rm -rf "$tmp"
