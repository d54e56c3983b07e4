# Compiles the reference's fibrin network code (network.cpp/network.h, read in
# place from REF) with ref_harness.cpp into oracle/_ref/libnetref.so.  Used only
# as a checker of the oracle.  The rest of the reference needs CUDA and GSL and
# is not built.
REF ?= /root/reference
CXX ?= g++

all: _ref/libnetref.so

_ref/libnetref.so: ref_harness.cpp $(REF)/network.cpp $(REF)/network.h
	@mkdir -p _ref
	$(CXX) -O2 -fPIC -shared -w -I$(REF) ref_harness.cpp $(REF)/network.cpp -o $@

.PHONY: all
