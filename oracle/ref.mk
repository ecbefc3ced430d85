# Builds the reference LMie C library from its own sources, where they lie, into
# oracle/_ref/ (git-ignored).  Test infrastructure only: the library is used as a
# checker and as the CPU baseline, never by the product path.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# The reference's own build system is not used: this compiles the C files of the
# path directly with gcc and the reference's stock flags (make.inc.example:12-20),
# with -DUSE_PTHREADS for its pthreads scheduler (src/Makefile:23-25).
# -Bsymbolic keeps the reference's calls to its own lmie_* symbols inside this
# library, so it can share a process with the product library that exports the
# same names.

REF      ?= /root/reference
OUT      := oracle/_ref
CC       := gcc
CFLAGS   := -O2 -fPIC -Wall -Wno-cpp -Wno-misleading-indentation -Wno-uninitialized \
            -Wno-unused-function -fno-strict-aliasing -fopenmp -DUSE_PTHREADS \
            -D'VERSION="0.91"' -I$(REF)/src -I$(REF)/misc
SRCS     := $(REF)/src/lmie_core.c $(REF)/src/lmie_interface.c $(REF)/src/lmie_schedule.c \
            $(REF)/misc/gutil_alloc.c $(REF)/misc/gutil_error.c $(REF)/misc/gutil_parse.c \
            $(REF)/misc/gindex_name_value.c $(REF)/misc/rtutil_math.c \
            $(REF)/misc/rtutil_scat_int.c $(REF)/misc/rtutil_size_dist.c \
            $(REF)/misc/rtutil_support.c $(REF)/misc/rtutil_scat_io.c $(REF)/misc/zeroin.c
OBJS     := $(patsubst $(REF)/%.c,$(OUT)/obj/%.o,$(SRCS))

# The reference's example program (examples/example_c.c), unmodified, compiled
# with the reference's own platform flags (common.inc:5) against this
# repository's include/lmie_interface.h and linked to the product library: the
# drop-in check of tests/test_gpu_fp64.py runs it on the GPU.
EXAMPLE  := $(OUT)/example_c_amd

# The reference's command line program (utils/calllmie.c), unmodified, linked
# statically with the objects above: the checker of the product's calllmie.
CALLLMIE := $(OUT)/calllmie_ref

all: $(OUT)/liblmie_ref.so $(EXAMPLE) $(CALLLMIE)

$(CALLLMIE): $(REF)/utils/calllmie.c $(OBJS)
	$(CC) $(CFLAGS) -o $@ $(REF)/utils/calllmie.c $(OBJS) -lpthread -lm

$(EXAMPLE): $(REF)/examples/example_c.c lmie_amd/lib/liblmie_amd.so
	@mkdir -p $(OUT)
	$(CC) -O2 -DPLATFORM=LINUX_GCC -DUNIX -Iinclude -o $@ $(REF)/examples/example_c.c -Llmie_amd/lib -llmie_amd \
	      -Wl,-rpath,'$$ORIGIN/../../lmie_amd/lib' -lm

$(OUT)/obj/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/liblmie_ref.so: $(OBJS)
	$(CC) -shared -fopenmp -Wl,-Bsymbolic -o $@ $(OBJS) -lpthread -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
