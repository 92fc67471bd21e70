# Compiles the reference's SYCL-free host code in place (nothing is copied):
#   - its unit tests test/test_nullable_semantics.cpp and test/test_sql_3vl.cpp
#   - oracle/ref_vectors.cpp, a harness of this repository that includes the
#     reference headers include/crystal/{pred,load,utils}.h and prints
#     known-answer vectors as JSON.
# Outputs go to oracle/_ref/ only (git-ignored).  join.h and inline_math.h
# need SYCL (absent), so they are not built.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CXX ?= g++
CXXFLAGS ?= -std=c++17 -O1

all: $(OUT)/test_nullable_semantics $(OUT)/test_sql_3vl $(OUT)/ref_vectors

$(OUT):
	mkdir -p $(OUT)

$(OUT)/test_nullable_semantics: $(REF)/test/test_nullable_semantics.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/include -o $@ $<

$(OUT)/test_sql_3vl: $(REF)/test/test_sql_3vl.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -o $@ $<

$(OUT)/ref_vectors: $(HERE)ref_vectors.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/include -o $@ $<

.PHONY: all
