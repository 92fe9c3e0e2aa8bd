#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/libvo_oracle.so.
set -e
cd "$(dirname "$0")"
g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared -Wall -o libvo_oracle.so vo_oracle.cpp -lpthread
