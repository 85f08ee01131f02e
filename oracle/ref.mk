# Compiles the reference's EOS and limiter sources where they lie (no copies),
# with the reference's own language level and no FMA contraction, and links
# them with ref_eos_driver.cpp into oracle/_ref/ref_eos.  The rest of the hot
# path (models, cells, relaxation) includes the MPI headers and is not built.
# Symbols of reference code that is never executed here (error reporting) are
# left unresolved at link time (the driver never calls it).
REF ?= /root/reference/src
OUT = _ref
CXX ?= g++
CXXFLAGS = -O3 -std=c++11 -ffp-contract=off -w -I$(REF)
SRCS = Eos/Eos.cpp Eos/EosSG.cpp Eos/EosIG.cpp Tools.cpp Maths/Coord.cpp Maths/Tensor.cpp \
       libTierces/tinyxml2.cpp Order2/Limiter.cpp Order2/LimiterMinmod.cpp Order2/LimiterVanLeer.cpp \
       Order2/LimiterVanAlbada.cpp Order2/LimiterSuperBee.cpp Order2/LimiterMC.cpp
OBJS = $(patsubst %.cpp,$(OUT)/%.o,$(SRCS))

all: $(OUT)/ref_eos

$(OUT)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/ref_eos: ref_eos_driver.cpp $(OBJS)
	$(CXX) $(CXXFLAGS) -no-pie -o $@ ref_eos_driver.cpp $(OBJS) -Wl,--unresolved-symbols=ignore-all

clean:
	rm -rf $(OUT)

.PHONY: all clean
