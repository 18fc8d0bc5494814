# Builds the reference CPU renderer and the fixture dumper from the reference
# sources where they lie (REF), with plain g++ -- no CMake, no OIDN, no GLFW.
# Outputs go only to oracle/_ref/ (git-ignored).  Usage:
#   make -f oracle/ref.mk            (from the repository root)
REF ?= /root/reference
SRC := $(REF)/code
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CXX ?= g++
# Same flags as the reference's own Linux build (ISO C++20, -O2, no FMA).
CXXFLAGS := -std=c++20 -O2 -DNDEBUG -DPCR_USE_GPU=0 -w -pthread -I$(SRC) -I$(SRC)/Includes
LIB_SRCS := NumGen.cpp CIE.cpp RGBToSpectrum.cpp SpdLoader.cpp LutDiscovery.cpp Bvh/Bvh.cpp \
            Scenes/SceneLoader.cpp Scenes/MeshLoader.cpp Scenes/SceneDiscovery.cpp \
            Scenes/Cornell.cpp Scenes/CornellSpheres.cpp Scenes/CornellLargeLight.cpp
CLI_SRCS := Cli/Main.cpp Renderer.cpp lodepng.cpp Denoise.cpp OidnDenoise.cpp
LIB_OBJS := $(patsubst %.cpp,$(OUT)/obj/%.o,$(LIB_SRCS))
CLI_OBJS := $(patsubst %.cpp,$(OUT)/obj/%.o,$(CLI_SRCS))

all: $(OUT)/pcr-ref-cli $(OUT)/refdump

$(OUT)/obj/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DPCR_BINARY_NAME='"frank-based-rendering-cli"' -c -o $@ $<

$(OUT)/pcr-ref-cli: $(LIB_OBJS) $(CLI_OBJS)
	$(CXX) $(CXXFLAGS) -o $@ $(filter %.cpp %.o,$^)

$(OUT)/refdump: $(HERE)../tests/golden/refdump.cpp $(HERE)../pcr_amd/host/procmesh.h $(LIB_OBJS)
	$(CXX) $(CXXFLAGS) -o $@ $(filter %.cpp %.o,$^)

clean:
	rm -rf $(OUT)

.PHONY: all clean
