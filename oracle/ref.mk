# Builds the reference's own headless LBM suites and its bench from the
# sources where they lie under /root/reference (nothing is copied into this
# repository).  Outputs go to oracle/_ref/ only, which git ignores.
#
#   make -f oracle/ref.mk            # all binaries
#   make -f oracle/ref.mk bench      # one of them
#
# Flags follow the reference's tests/Makefile (-std=c++17 -O1 -fopenmp, no
# -march): the baseline x86-64 target has no FMA, so every float operation is
# rounded on its own.  The reference's published bench checksum
# (rho 414728.371435) comes from a -march=native build where GCC contracts
# a*b+c into FMA (C++ dialects default to -ffp-contract=fast); bench_fma
# reproduces that number, bench is the uncontracted arithmetic that the
# oracle and the HIP kernels follow.
REF     ?= /root/reference/aeros
HERE    := $(dir $(lastword $(MAKEFILE_LIST)))
OUT     := $(HERE)_ref
SRC     := $(REF)/src
T       := $(REF)/tests
INC     := -I$(SRC) -I$(REF)/../libs/glm -I$(REF)/../libs/glad/include -I$(REF)/../libs/glfw/include
CXX     ?= g++
FLAGS   := -std=c++17 -O1 -fopenmp -w $(INC)
CORE    := $(SRC)/lbm.cpp $(SRC)/thermal_lbm.cpp $(SRC)/flow_field.cpp $(SRC)/voxel_grid.cpp \
           $(SRC)/atmosphere.cpp $(SRC)/globals.cpp $(SRC)/debug_log.cpp
COMMON  := $(SRC)/globals.cpp $(SRC)/debug_log.cpp

BINS := bench bench_fma lbm physref thermal validate voxel ref_fields ref_box ref_voxels
.PHONY: all $(BINS)
all: $(BINS)

$(OUT):
	mkdir -p $(OUT)

bench: | $(OUT)
	$(CXX) -std=c++17 -O3 -fopenmp -w $(INC) $(T)/bench_lbm.cpp $(CORE) -o $(OUT)/bench -pthread

bench_fma: | $(OUT)
	$(CXX) -std=c++17 -O3 -march=native -fopenmp -w $(INC) $(T)/bench_lbm.cpp $(CORE) -o $(OUT)/bench_fma -pthread

lbm: | $(OUT)
	$(CXX) $(FLAGS) $(T)/lbm_physics_headless.cpp $(CORE) -o $(OUT)/lbm -pthread

physref: | $(OUT)
	$(CXX) $(FLAGS) $(T)/physics_ref_headless.cpp $(CORE) -o $(OUT)/physref -pthread

thermal: | $(OUT)
	$(CXX) $(FLAGS) $(T)/thermal_headless.cpp $(CORE) -o $(OUT)/thermal -pthread

validate: | $(OUT)
	$(CXX) $(FLAGS) $(T)/validation_headless.cpp $(CORE) -o $(OUT)/validate -pthread

voxel: | $(OUT)
	$(CXX) $(FLAGS) $(T)/voxel_headless.cpp $(SRC)/voxel_grid.cpp $(COMMON) -o $(OUT)/voxel -pthread

# Per-cell field dumps of the reference solver (oracle/ref_fields.cpp, test
# infrastructure); the link stubs come from the reference bench's own
# translation unit with its main renamed.
ref_fields: | $(OUT)
	$(CXX) $(FLAGS) -Dmain=aeros_bench_main_unused -c $(T)/bench_lbm.cpp -o $(OUT)/bench_stubs.o
	$(CXX) $(FLAGS) $(HERE)ref_fields.cpp $(OUT)/bench_stubs.o $(CORE) -o $(OUT)/ref_fields -pthread

# The same harness built with the reference bench's flags (-O3), for timing
# the reference solver on the measured case (bench.py's CPU baseline).
ref_box: | $(OUT)
	$(CXX) -std=c++17 -O3 -fopenmp -w $(INC) -Dmain=aeros_bench_main_unused -c $(T)/bench_lbm.cpp -o $(OUT)/bench_stubs_o3.o
	$(CXX) -std=c++17 -O3 -fopenmp -w $(INC) $(HERE)ref_fields.cpp $(OUT)/bench_stubs_o3.o $(CORE) -o $(OUT)/ref_box -pthread

# The reference's voxel vote over a given grid (oracle/ref_voxels.cpp, test
# infrastructure): header-only BVH, no link stubs needed.
ref_voxels: | $(OUT)
	$(CXX) $(FLAGS) $(HERE)ref_voxels.cpp -o $(OUT)/ref_voxels
