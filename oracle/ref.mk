# Builds the reference's own FP64 CPU engine (yeatanir/Metal-Reax "cpu64") from its sources where they lie, so that the
# oracle restatement in this directory can be checked against it. Outputs go only to oracle/_ref/ (git-ignored).
# Usage: make -f oracle/ref.mk REF=/root/reference        (run from the repository root)
# The reference is compiled with its own strict-FP flags (-ffp-contract=off -fno-fast-math, cmake/ReaxMetalFlags.cmake:16-22).
REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
CXXFLAGS := -std=c++20 -O2 -fPIC -ffp-contract=off -fno-fast-math -I$(REF)/include -I$(REF)/src/metal/shaders -w
SRCS := src/core/capabilities.cpp src/core/compat_flags.cpp src/core/energy_terms.cpp src/core/pins.cpp src/core/sha256.cpp \
        src/core/system.cpp src/cpu/bonded.cpp src/cpu/nonbonded.cpp src/cpu/bonded_fp32.cpp src/cpu/nonbonded_fp32.cpp src/cpu/acks2.cpp src/neighbor/neighbor.cpp \
        src/neighbor/nonbonded_device.cpp src/neighbor/qeq_device.cpp src/neighbor/bonded_device.cpp src/io/ffield_parser.cpp \
        src/io/control_parser.cpp src/io/pair_settings.cpp src/io/table_dump.cpp src/metal/metal_backend_stub.cpp
OBJS := $(patsubst %.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/reaxmetal_energy_tool $(OUT)/reaxmetal_ffield_dump $(OUT)/ref_peratom

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/reaxmetal_energy_tool: $(OBJS) $(REF)/tools/energy_tool.cpp
	$(CXX) $(CXXFLAGS) $(REF)/tools/energy_tool.cpp $(OBJS) -o $@ -lpthread

# per-atom tallies through the reference's public interface (a driver of this directory, not a reference source)
$(OUT)/ref_peratom: $(OBJS) oracle/ref_peratom.cpp
	$(CXX) $(CXXFLAGS) oracle/ref_peratom.cpp $(OBJS) -o $@ -lpthread

$(OUT)/reaxmetal_ffield_dump: $(OBJS) $(REF)/tools/ffield_dump.cpp
	$(CXX) $(CXXFLAGS) $(REF)/tools/ffield_dump.cpp $(OBJS) -o $@ -lpthread

clean:
	rm -rf $(OUT)
.PHONY: all clean
