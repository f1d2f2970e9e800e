# Builds the reference's own C++ search path from its sources where they lie
# (read-only) into oracle/_ref/.  Test infrastructure only: the modules built
# here are used to write golden vectors and as the timed CPU baseline, never
# as part of the product.  Same flags as the reference's setup.py:10 except
# -march: x86-64-v4 (AVX-512 F/BW/DQ/VL) instead of native, so the objects run
# on any AVX-512 host and take the same AVX-512 branch of simd.h (src/simd.h:3-11).
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
PY_INC   := $(shell python3 -c "import sysconfig;print(sysconfig.get_paths()['include'])")
PB_INC   := $(shell python3 -c "import pybind11;print(pybind11.get_include())")
EXT      := $(shell python3 -c "import sysconfig;print(sysconfig.get_config_var('EXT_SUFFIX'))")
CXX      ?= g++
REFFLAGS := -O3 -ffast-math -march=x86-64-v4 -mtune=generic -pthread -fPIC -shared -fvisibility=hidden

all: $(OUT)/linear_ann_cpp$(EXT) $(OUT)/ivf_lsh_cpp$(EXT) $(OUT)/hnsw_cpp$(EXT)

# src/linear_index.cpp + src/simd.h + src/thread_pool.h  (setup.py:5-12)
$(OUT)/linear_ann_cpp$(EXT): $(REF)/src/linear_index.cpp $(REF)/src/simd.h
	@mkdir -p $(OUT)
	$(CXX) -std=c++14 $(REFFLAGS) -I$(REF)/src -I$(PY_INC) -I$(PB_INC) $< -o $@

# competitors/ivf_lsh/src/lsh_index_optimized.cpp with ITS OWN simd.h (setup.py:50-57)
$(OUT)/ivf_lsh_cpp$(EXT): $(REF)/competitors/ivf_lsh/src/lsh_index_optimized.cpp $(REF)/competitors/ivf_lsh/src/simd.h
	@mkdir -p $(OUT)
	$(CXX) -std=c++14 $(REFFLAGS) -I$(REF)/competitors/ivf_lsh/src -I$(PY_INC) -I$(PB_INC) $< -o $@

.PHONY: all

# submission/src/hnsw.cpp with ITS OWN simd.h (submission/setup.py:4-13): the
# reference's HNSW (per-dim SQ8, heuristic neighbours, reorder), timed beside
# the GPU path as a CPU baseline
$(OUT)/hnsw_cpp$(EXT): $(REF)/submission/src/hnsw.cpp $(REF)/submission/src/simd.h
	@mkdir -p $(OUT)
	$(CXX) -std=c++17 $(REFFLAGS) -I$(REF)/submission/src -I$(PY_INC) -I$(PB_INC) $< -o $@
