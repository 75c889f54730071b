# Builds the reference HiParTI C sources, where they lie under
# /root/reference, into oracle/_ref/libhiparti_ref.so.  Used only in the build
# container to pin the oracle (tests/golden/make_goldens.py); the library is
# never shipped or loaded by the product.
#
# Flags follow the reference's OpenMP release configuration
# (CMakeLists.txt:85-90: -D_GNU_SOURCE -O3 -fopenmp -DADD_, OpenMP on).
# The four files that call BLAS/LAPACK (ssyrk_, spotrf_, spotrs_, sgesv_),
# which this image does not provide, are left out: the CPD-ALS drivers and the
# normal-equation solvers.  Everything on the MTTKRP and HiCOO-conversion path
# builds and links with no undefined symbols.
#
#   make -f oracle/ref.mk            (from the repository root)

REF      ?= /root/reference
OUT      := oracle/_ref
LIB      := $(OUT)/libhiparti_ref.so
CC       ?= gcc
CFLAGS   := -std=gnu99 -D_GNU_SOURCE -O3 -fopenmp -DADD_ -DHIPARTI_USE_OPENMP -DNDEBUG -fPIC \
            -DHIPARTI_INDEX_TYPEWIDTH=32 -DHIPARTI_VALUE_TYPEWIDTH=32 -DHIPARTI_ELEMENT_INDEX_TYPEWIDTH=8 \
            -w -I$(REF)/include

LAPACK_USERS := $(REF)/src/matrix/matrix_solver.c $(REF)/src/matrix/rankmatrix_solver.c \
                $(REF)/src/sptensor/cpd.c $(REF)/src/sptensor/cpd_omp.c \
                $(REF)/src/sptensor/hicoo/cpd.c $(REF)/src/sptensor/hicoo/cpd_omp.c
SRCS := $(filter-out $(LAPACK_USERS),$(shell find $(REF)/src -name '*.c' | sort))
OBJS := $(patsubst $(REF)/src/%.c,$(OUT)/obj/%.o,$(SRCS))

all: $(LIB)

$(OUT)/obj/%.o: $(REF)/src/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(LIB): $(OBJS)
	$(CC) -shared -fopenmp -Wl,--no-undefined -o $@ $(OBJS) -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
