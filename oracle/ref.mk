# Builds checkers from the reference's own sources, where they lie, into
# oracle/_ref/ (git-ignored, but kept with the working tree so the built
# checkers are available wherever the tree is copied).  Nothing is copied from
# the reference.  Needs the reference tree at /root/reference.
#   _ref/libref_gridfp.so : the reference's src/common/gridfp_transition.hpp and
#                           gridfp_reverse.hpp behind ref_gridfp_harness.cpp
#   _ref/ggcount_single   : third_party/ggcount/src/ggcount.cpp (MIT), an
#                           independent CPU frontier DP (Iwashita et al. 2013)
#   _ref/ggcount_omp      : the same source built with its own OpenMP loops
#                           (used for the n=20..22 fixtures)
REF ?= /root/reference
OUT := _ref

all: $(OUT)/libref_gridfp.so $(OUT)/ggcount_single $(OUT)/ggcount_omp

$(OUT):
	mkdir -p $(OUT)

$(OUT)/libref_gridfp.so: ref_gridfp_harness.cpp $(REF)/src/common/gridfp_transition.hpp $(REF)/src/common/gridfp_reverse.hpp | $(OUT)
	g++ -std=c++17 -O2 -fPIC -shared -I$(REF)/src/common ref_gridfp_harness.cpp -o $@

$(OUT)/ggcount_single: $(REF)/third_party/ggcount/src/ggcount.cpp | $(OUT)
	g++ -std=c++11 -DNDEBUG -O3 $< -o $@

$(OUT)/ggcount_omp: $(REF)/third_party/ggcount/src/ggcount.cpp | $(OUT)
	g++ -std=c++11 -DNDEBUG -O3 -fopenmp $< -o $@

.PHONY: all
