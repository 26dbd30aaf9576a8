# Builds the reference's graph_mining_system from its own source files, in place,
# with gcc/g++ only (no cmake, no configure).  Every header it needs is a file the
# reference ships: nauty.h / gtools.h (utility/automorphism), GLPK's config.h and
# src/*.h (utility/td/glpk-4.55).  Nothing is written outside oracle/_ref/.
#
#   make -f oracle/ref.mk -j8           (from the repo root)
#
# The binary is the CPU checker and the "reference" CPU baseline; it is never the
# product path.  -march=native is replaced by -mavx2 -mfma (the reference selects
# its AVX2 kernels with SI=0, config/config.h:27-31) so the binary also runs on a
# different x86 host.

REF     ?= /root/reference
OUT     ?= oracle/_ref
OBJ     := $(OUT)/obj
CC      ?= gcc
CXX     ?= g++
ARCHF   := -mavx2 -mfma -mpopcnt -mbmi -mbmi2
CFLAGS_N := -O3 -fPIC -w
CFLAGS_G := -O2 -fPIC -w -DHAVE_CONFIG_H
CXXFLAGS := -O3 -std=c++17 -fopenmp $(ARCHF) -w

GLPK    := $(REF)/utility/td/glpk-4.55
GLPKSRC := $(GLPK)/src
GLPK_C  := $(shell sed -n '/libglpk_la_SOURCES/,/^$$/p' $(GLPKSRC)/Makefile.am | grep '\.c' | sed 's/ *\\//;s/^ *//')
GLPK_INC := -I$(GLPK) -I$(GLPKSRC) $(addprefix -I$(GLPKSRC)/,amd bflib cglib colamd env minisat misc proxy zlib)

NAUTY   := $(REF)/utility/automorphism
NAUTY_C := nauty.c nautil.c nausparse.c naugraph.c schreier.c naurng.c naugroup.c gtools.c gutil1.c gutil2.c nautinv.c

REF_CXX := main.cpp \
  matching/join.cpp matching/parallel.cpp \
  relation/candidate_space.cpp relation/labeled_graph.cpp relation/relation.cpp relation/graph.cpp \
  plan/decomposition.cpp plan/estimator.cpp plan/optimizer.cpp plan/subset_structure.cpp plan/all_td.cpp \
  utility/command_parser.cpp utility/utils.cpp utility/command.cpp utility/compute_set_intersection.cpp utility/clique.cpp \
  utility/td/dataloader.cpp utility/td/hypergraph.cpp utility/td/preprocessing.cpp utility/td/td.cpp \
  utility/td/twbound.cpp utility/td/SCsolver.cpp

REF_INC := -I$(REF)/config -I$(REF)/utility -I$(REF)/relation -I$(REF)/plan -I$(REF)/matching \
  -I$(REF)/utility/automorphism -I$(REF)/utility/td -I$(REF)/utility/td/glpk_lib/include

GLPK_O  := $(addprefix $(OBJ)/glpk/,$(GLPK_C:.c=.o))
NAUTY_O := $(addprefix $(OBJ)/nauty/,$(NAUTY_C:.c=.o))
REF_O   := $(addprefix $(OBJ)/ref/,$(REF_CXX:.cpp=.o))

all: $(OUT)/graph_mining_system

$(OUT)/graph_mining_system: $(REF_O) $(NAUTY_O) $(GLPK_O)
	$(CXX) -fopenmp -o $@ $^ -lm

$(OBJ)/glpk/%.o: $(GLPKSRC)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS_G) $(GLPK_INC) -c $< -o $@

$(OBJ)/nauty/%.o: $(NAUTY)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS_N) -I$(NAUTY) -c $< -o $@

$(OBJ)/ref/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(REF_INC) -c $< -o $@

.PHONY: all
