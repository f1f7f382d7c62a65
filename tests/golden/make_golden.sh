#!/bin/sh
# Regenerates the libstdc++ uniform_int golden vectors with this image's g++.
set -e
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
g++ -O2 -std=c++17 -o "$tmp/gen" "$here/gen_libstdcxx_uniform.cpp"
"$tmp/gen" > "$here/libstdcxx_uniform.json"
rm -rf "$tmp"
