# Drop-in LPMD plugins over the MI355X engine, built against the reference's
# headers and core library.
#
#   make -f lpmd2_amd/plugins/plugins.mk       (from the repository root)
#
# Needs the reference headers ($(REF)/liblpmd) and its core library to link
# against: here the core built from the reference sources by oracle/ref.mk
# (oracle/_ref), in a real deployment the reference's own liblpmd (CORE=...).
# The plugins carry no run path to the core: they use the one the host
# program already loaded.  Writes lpmd2_amd/plugins/_build/{liblpmdx_lpmd.so,
# <name>.so}: putting that directory first on LPMD_PATH replaces the CPU
# plugins of the same names.
REF      ?= /root/reference
OUT      := lpmd2_amd/plugins/_build
CORE     ?= oracle/_ref
CORELIB  ?= lpmdref_core
CXX      ?= g++
CXXFLAGS := -O2 -std=c++17 -fPIC -w -I$(REF)/liblpmd -Iinclude
NAMES    := lennardjones fastlj morse buckingham suttonchen finnissinclair gupta \
            harmonic lennardjonesMod finnissinclair-ext tabulatedpair \
            linkedcell verletlist minimumimage lcbinary velocityverlet leapfrog beeman nosehoover
RPATH    := -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../..'

all: $(patsubst %,$(OUT)/%.so,$(NAMES)) $(OUT)/libsim_driver.so

$(OUT)/liblpmdx_lpmd.so: lpmd2_amd/plugins/lpmdx_lpmd.cc lpmd2_amd/plugins/lpmdx_lpmd.hpp include/lpmdx.h \
                         lpmd2_amd/csrc/table.hpp \
                         lpmd2_amd/liblpmdx.so $(CORE)/liblpmdref_core.so
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $< -L$(CORE) -l$(CORELIB) -Llpmd2_amd -l:liblpmdx.so $(RPATH) \
	  -Wl,--no-undefined

$(OUT)/%.so: lpmd2_amd/plugins/plugin_entry.cc $(OUT)/liblpmdx_lpmd.so
	$(CXX) $(CXXFLAGS) -DLPMDX_PLUGIN_NAME='"$*"' -shared -o $@ $< -L$(OUT) -l:liblpmdx_lpmd.so \
	  -L$(CORE) -l$(CORELIB) $(RPATH) -Wl,--no-undefined

# test driver of DeviceSimulation (tests only: C entry points for ctypes)
$(OUT)/libsim_driver.so: tests/dropin_driver/sim_driver.cc $(OUT)/liblpmdx_lpmd.so
	$(CXX) $(CXXFLAGS) -shared -o $@ $< -L$(OUT) -l:liblpmdx_lpmd.so -L$(CORE) -l$(CORELIB) $(RPATH) -ldl \
	  -Wl,--no-undefined

clean:
	rm -rf $(OUT)

.PHONY: all clean
