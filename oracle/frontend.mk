# Builds the reference's own library (front end, scheduler, interpreter and
# its OpenCL translator) from its sources where they lie, into
# oracle/_ref/libchai_frontend.a, for this machine's tests only: its
# interpreter is the checker the restated oracle and the GPU library are
# compared with, and its scheduler hosts the MI355X device class
# (chai_amd/device/) in the integration tests.  Nothing of the reference is
# copied: objects are compiled from /root/reference, and the `chai/` include
# directory the reference's sources expect is a directory of symbolic links
# under oracle/_ref/include.
#
# Usage: make -C oracle -f frontend.mk [-j8]
REFSRC  ?= /root/reference
CXX     ?= g++
OUT     := _ref/frontend
LIBDIRS := misc Random123 vendor autotune kernel bytecode common memmgr interp \
           jitast jitbc jitstmt jitkrnl enqueue runtime api applib
# every .cpp of those directories except jitkrnl/ScalarArgs.cpp, which the
# reference's own jitkrnl/Makefile leaves out (it no longer compiles)
SRCS    := $(filter-out %/jitkrnl/ScalarArgs.cpp,$(foreach d,$(LIBDIRS),$(wildcard $(REFSRC)/$(d)/*.cpp)))
OBJS    := $(foreach s,$(SRCS),$(OUT)/$(notdir $(patsubst %/,%,$(dir $(s))))__$(notdir $(s:.cpp=.o)))
INCS    := -I_ref/include $(addprefix -I$(REFSRC)/,$(LIBDIRS))
# the reference's own flags (Makefile.common, vendor/Makefile.AMD) plus what
# a current g++ needs for this 2012 C++98 code base
REFFLAGS := -O2 -std=gnu++98 -w -include cstddef -fopenmp -D_VENDOR_IS_AMD_ \
            -DCL_USE_DEPRECATED_OPENCL_1_1_APIS -DCL_TARGET_OPENCL_VERSION=120

# public headers the sources include as <chai/...> (installed there by the
# reference's build: api/, bytecode/, applib/, extensions/ Makefiles)
PUBHDRS := api/chai.h bytecode/BC.hpp bytecode/RefCnt.hpp bytecode/Stak.hpp bytecode/Visit.hpp \
           applib/ParseArgs.hpp extensions/chaiExt.h

all: _ref/libchai_frontend.a

_ref/include/chai/.stamp:
	mkdir -p _ref/include/chai
	for h in $(PUBHDRS); do ln -sf $(REFSRC)/$$h _ref/include/chai/$$(basename $$h); done
	touch $@

define objrule
$(OUT)/$(1)__%.o: $(REFSRC)/$(1)/%.cpp _ref/include/chai/.stamp
	@mkdir -p $(OUT)
	$$(CXX) $$(REFFLAGS) $$(INCS) -c $$< -o $$@
endef
$(foreach d,$(LIBDIRS),$(eval $(call objrule,$(d))))

# api/ArrayBase.cpp has two non-void functions that end without a return
# (lines 129, 143).  Current g++ treats that end as unreachable at -O1 and
# above, which corrupts the caller; at -O0 it returns, as the 2012 compilers
# did.
$(OUT)/api__ArrayBase.o: REFFLAGS := $(subst -O2,-O0,$(REFFLAGS))

_ref/libchai_frontend.a: $(OBJS)
	rm -f $@
	ar qc $@ $(OBJS)
	ranlib $@

clean:
	rm -rf $(OUT) _ref/libchai_frontend.a _ref/include

.PHONY: all clean
