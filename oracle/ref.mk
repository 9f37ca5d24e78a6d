# Builds the reference Fortran library (flatironinstitute/fmm2d) from its own
# source files, where they lie under REFROOT, into oracle/_ref/.  Test
# infrastructure only: the product never links or loads this library.
#
# Compiler: ROCm LLVM flang (gfortran, the reference's default, is absent).
# common/cumsum.f is compiled without -fopenmp because flang rejects the
# directive spelling at common/cumsum.f:216.
REFROOT ?= /root/reference
FLANG   ?= /opt/rocm/lib/llvm/bin/flang
OUT     := $(dir $(lastword $(MAKEFILE_LIST)))_ref
FFLAGS  := -fPIC -O3 -march=native -w -fopenmp
SRC     := $(REFROOT)/src

COMMON := cdjseval2d dfft_threadsafe fmmcommon2d next235 prini tree_routs2d \
          pts_tree2d hank103 hkrand dlaran
LAP    := cauchykernels2d cfmm2d cfmm2d_ndiv cfmm2dwrap cfmm2dwrap_vec l2dterms \
          lapkernels2d laprouts2d lfmm2d lfmm2d_ndiv lfmm2dwrap lfmm2dwrap_vec \
          lndiv2d rfmm2d rfmm2d_ndiv rfmm2dwrap rfmm2dwrap_vec rlapkernels2d
HELM   := h2dcommon h2dterms helmkernels2d helmrouts2d hfmm2d hfmm2d_ndiv \
          hfmm2dwrap hfmm2dwrap_vec hndiv2d wideband2d hfmm2d_mps hfmm2d_mps_ndiv
BH     := bh2dterms bhfmm2d bhfmm2dwrap bhkernels2d bhndiv2d bhrouts2d
ST     := stfmm2d stokkernels2d

OBJS := $(addprefix $(OUT)/obj/c_,$(addsuffix .o,$(COMMON))) \
        $(OUT)/obj/c_cumsum.o \
        $(addprefix $(OUT)/obj/l_,$(addsuffix .o,$(LAP))) \
        $(addprefix $(OUT)/obj/h_,$(addsuffix .o,$(HELM))) \
        $(addprefix $(OUT)/obj/b_,$(addsuffix .o,$(BH))) \
        $(addprefix $(OUT)/obj/s_,$(addsuffix .o,$(ST)))

all: $(OUT)/libfmm2d_ref.so

$(OUT)/obj:
	mkdir -p $@

$(OUT)/obj/c_cumsum.o: $(SRC)/common/cumsum.f | $(OUT)/obj
	$(FLANG) -fPIC -O3 -march=native -w -c $< -o $@
$(OUT)/obj/c_%.o: $(SRC)/common/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@
$(OUT)/obj/l_%.o: $(SRC)/laplace/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@
$(OUT)/obj/h_%.o: $(SRC)/helmholtz/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@
$(OUT)/obj/b_%.o: $(SRC)/biharmonic/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@
$(OUT)/obj/s_%.o: $(SRC)/stokes/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/libfmm2d_ref.so: $(OBJS)
	$(FLANG) -shared -fopenmp -o $@ $^ -Wl,-rpath,/opt/rocm/lib/llvm/lib

clean:
	rm -rf $(OUT)
.PHONY: all clean
