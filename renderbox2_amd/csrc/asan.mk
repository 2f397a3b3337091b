# Sanitizer build of librb.so for the CPU test suite: `make -f asan.mk asan-test`
# (this container only; GPU hosts run the release build).
include Makefile

# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build: the host
# sources (scene API, builder, layouts, BVH cache, OBJ/XML readers, image
# writers) are instrumented, the HIP sources are the release objects.  The
# library leaves the sanitizer runtime to the process, so run it with gcc's
# runtime preloaded (CPU only: `make -f asan.mk asan-test`).
ASAN_OUT ?= ../../_build/librb_asan.so
ASAN_BUILD ?= ../../_build/rb_asan
ASAN_FLAGS = -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined
ASAN_OBJ = $(patsubst %.cpp,$(ASAN_BUILD)/%.o,$(CPP_SRC))

$(ASAN_BUILD):
	mkdir -p $(ASAN_BUILD)

$(ASAN_BUILD)/%.o: %.cpp $(HDRS) | $(ASAN_BUILD)
	$(CXX) $(CXXFLAGS) $(ASAN_FLAGS) -c $< -o $@

$(ASAN_OUT): $(ASAN_OBJ) $(HIP_OBJ)
	$(CXX) -shared -fPIC -o $@.tmp $(ASAN_OBJ) $(HIP_OBJ) -L/opt/rocm/lib -lamdhip64 -pthread \
	    -Wl,-rpath,/opt/rocm/lib && mv $@.tmp $@

asan: $(ASAN_OUT)

# The CPU test suite against the sanitizer build (leak checks off: the
# interpreter itself keeps allocations to exit).
asan-test: $(ASAN_OUT)
	cd ../.. && RB_LIB=$(abspath $(ASAN_OUT)) ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 \
	    UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
	    LD_PRELOAD="$$($(CXX) -print-file-name=libasan.so) $$($(CXX) -print-file-name=libubsan.so)" \
	    python -m pytest tests -x -q -m "not gpu" -p no:cacheprovider

asan-clean:
	rm -rf $(ASAN_BUILD) $(ASAN_OUT)

.PHONY: asan asan-test asan-clean
