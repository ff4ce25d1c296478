# Builds the reference TAL-SH CPU path (CP-TAL + talsh.h C API, -DNO_GPU -DNO_BLAS)
# from the sources where they lie under /root/reference/TALSH, into oracle/_ref/.
# Test infrastructure only: the product never links or loads this library.
#   make -f oracle/ref.mk            (from the repository root)
REF   ?= /root/reference/TALSH
OUT   ?= oracle/_ref
FC    := /opt/rocm/lib/llvm/bin/amdflang
CXX   ?= g++
DEFS   = -DNO_GPU -DNO_AMD -DNO_PHI -DNO_BLAS -DLINUX
FFLAGS = -c -fopenmp -O3 -fPIC -cpp $(DEFS) -J $(OUT)/mod
CFLAGS = -c -fopenmp -O3 -fPIC -std=c++11 $(DEFS) -w
FSRC   = dil_basic stsubs combinatoric symm_index timers tensor_algebra tensor_algebra_cpu tensor_algebra_cpu_phi talshf
CSRC   = timer byte_packet mem_manager tensor_algebra_gpu talshc
FOBJ   = $(addprefix $(OUT)/obj/,$(addsuffix .o,$(FSRC)))
COBJ   = $(addprefix $(OUT)/obj/,$(addsuffix .o,$(CSRC)))
FRT    = /opt/rocm/lib/llvm/lib/clang/22/lib/x86_64-unknown-linux-gnu
LLIB   = /opt/rocm/lib/llvm/lib

all: $(OUT)/libtalsh_ref.so

$(OUT)/obj $(OUT)/mod:
	mkdir -p $@

# Fortran modules must be compiled in dependency order (serial), hence one stamp rule.
$(OUT)/obj/fortran.stamp: | $(OUT)/obj $(OUT)/mod
	set -e; for f in $(FSRC); do $(FC) $(FFLAGS) $(REF)/$$f.F90 -o $(OUT)/obj/$$f.o; done
	touch $@

$(OUT)/obj/%.o: $(REF)/%.cpp | $(OUT)/obj
	$(CXX) $(CFLAGS) $< -o $@

$(OUT)/libtalsh_ref.so: $(OUT)/obj/fortran.stamp $(COBJ)
	$(CXX) -shared -o $@ $(FOBJ) $(COBJ) -L$(FRT) -Wl,--whole-archive -lflang_rt.runtime -Wl,--no-whole-archive \
	  -L$(LLIB) -Wl,-rpath,$(LLIB) -lomp -latomic -lm

clean:
	rm -rf $(OUT)
