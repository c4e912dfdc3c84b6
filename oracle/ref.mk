# Builds the reference CAVIAR library from its own sources, where they lie
# under /root/reference, plus the golden-vector harness that links against it.
# Outputs go only to oracle/_ref/.  The reference's CMake build is not used:
# every source under src/caviar/ is compiled directly with g++ (serial build,
# all optional dependencies off, which is the reference's default).
#
#   make -f oracle/ref.mk -j8          (from the repository root)
REF      ?= /root/reference
OUT      := oracle/_ref
CXX      ?= g++
CXXFLAGS := -std=c++17 -O3 -w -I$(REF)/include

SRCS := $(shell find $(REF)/src/caviar -name '*.cpp' | sort)
OBJS := $(patsubst $(REF)/src/caviar/%.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/ref_harness

$(OUT)/obj/%.o: $(REF)/src/caviar/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libcaviar_ref.a: $(OBJS)
	rm -f $@
	ar rcs $@ $(OBJS)

$(OUT)/ref_harness: oracle/ref_harness.cpp $(OUT)/libcaviar_ref.a
	$(CXX) $(CXXFLAGS) oracle/ref_harness.cpp $(OUT)/libcaviar_ref.a -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
