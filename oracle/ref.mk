# Partial build of the LPMD reference, used only as a checker.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Compiles the reference's own sources where they lie under $(REF) with g++
# directly (the reference's CMake build is not used) and writes every output
# into oracle/_ref/.  The session-bound pieces (liblpmd/runtime/session.cc,
# liblpmd/plugins/pluginmanager.cc) include the CMake-generated config.h and
# are therefore left out; nothing below needs them.
REF      ?= /root/reference
OUT      := oracle/_ref
CXX      ?= g++
# -O2 without -march: the reference's Release flags on x86-64 (no FMA contraction)
CXXFLAGS := -O2 -std=c++17 -fPIC -w -I$(REF)/liblpmd -I$(REF)/plugins

CORE_SRCS := \
  liblpmd/core/atom.cc liblpmd/core/cell.cc liblpmd/core/cellmanager.cc \
  liblpmd/core/configuration.cc liblpmd/core/matrix.cc liblpmd/core/paramlist.cc \
  liblpmd/core/particleset.cc liblpmd/core/value.cc \
  liblpmd/plugins/module.cc liblpmd/plugins/moduleinfo.cc liblpmd/plugins/plugin.cc \
  liblpmd/support/error.cc liblpmd/support/util.cc liblpmd/support/elements.cc liblpmd/support/colorhandler.cc \
  liblpmd/simulation/potential.cc liblpmd/simulation/pairpotential.cc \
  liblpmd/simulation/metalpotential.cc liblpmd/simulation/cellgenerator.cc \
  liblpmd/simulation/integrator.cc liblpmd/simulation/twostepintegrator.cc \
  liblpmd/simulation/onestepintegrator.cc \
  liblpmd/simulation/simulation.cc liblpmd/simulation/systemmodifier.cc

PLUGINS := linkedcell verletlist minimumimage lcbinary lennardjones fastlj morse buckingham tabulatedpair \
           suttonchen finnissinclair gupta harmonic lennardjonesMod finnissinclair-ext \
           crystal3d velocityverlet leapfrog beeman

CORE_OBJS := $(patsubst %.cc,$(OUT)/obj/%.o,$(CORE_SRCS))
PLUGIN_SOS := $(patsubst %,$(OUT)/%.so,$(PLUGINS))

all: $(OUT)/libref_harness.so $(PLUGIN_SOS)

$(OUT)/obj/%.o: $(REF)/%.cc
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/liblpmdref_core.so: $(CORE_OBJS)
	$(CXX) -shared -o $@ $^ -ldl -Wl,--no-undefined

$(OUT)/%.so: $(REF)/plugins/%.cc $(OUT)/liblpmdref_core.so
	$(CXX) $(CXXFLAGS) -shared -o $@ $< -L$(OUT) -llpmdref_core -Wl,-rpath,'$$ORIGIN' -Wl,--no-undefined

$(OUT)/libref_harness.so: oracle/ref_harness/ref_harness.cc $(OUT)/liblpmdref_core.so $(OUT)/verletlist.so
	$(CXX) $(CXXFLAGS) -shared -o $@ $< -L$(OUT) -llpmdref_core -l:verletlist.so -Wl,-rpath,'$$ORIGIN' -ldl -Wl,--no-undefined

clean:
	rm -rf $(OUT)

.PHONY: all clean
