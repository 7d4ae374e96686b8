# Builds the reference's own C++ sources that compile here, as checkers for
# the product's restatements -- TEST INFRASTRUCTURE ONLY.  Sources are read
# where they lie under the reference checkout; outputs go to oracle/_ref/.
#
#   libsky_ref.so: host/sky/sky_nishita.cpp (Blender's Nishita sky
#   precompute, with its own sky_float3.h and sky_model.h), exporting
#   SKY_nishita_skymodel_precompute_texture / _sun.
REFERENCE ?= /root/reference
SKY = $(REFERENCE)/raymond/host/sky
CXX ?= g++
CXXFLAGS = -O2 -std=c++17 -fPIC -ffp-contract=off -Wno-unused-function

all: _ref/libsky_ref.so

_ref/libsky_ref.so: $(SKY)/sky_nishita.cpp $(SKY)/sky_float3.h $(SKY)/sky_model.h
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -I$(SKY) -o $@ $(SKY)/sky_nishita.cpp

.PHONY: all
