# Builds the reference's own CPU PDHG path (no CUDA: the reference's no-GPU stub
# stands in for its GPU backend) straight from the sources under REF, with g++.
# Everything it writes goes under oracle/_ref/, which git ignores.  The outputs
# are a test checker and a CPU timing baseline, never part of the product.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
REF ?= /root/reference
OUT := oracle/_ref
OBJ := $(OUT)/obj
SRCS := model certificate io verify preprocess bound_postsolve cpu solver mip research \
        simplex cuts barrier portfolio advisor sensitivity iis global history decompose gnn \
        cuda_stub
CXX := g++
CXXFLAGS := -std=c++20 -O2 -fopenmp -fPIC -w -I$(REF)/include -I$(REF)/third_party \
            -isystem $(REF)/third_party/eigen -DVANTAGE_SPARSE_LU -DEIGEN_DONT_PARALLELIZE \
            -DNIRYUKTI_VERSION=\"0.2.2\"
OBJS := $(addprefix $(OBJ)/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/libvantage_ref.so $(OUT)/niryukti_ref $(OUT)/ref_driver

$(OBJ)/%.o: $(REF)/src/%.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ):
	mkdir -p $(OBJ)

$(OUT)/libvantage_ref.so: $(OBJS) $(OBJ)/c_api.o
	$(CXX) -shared -fopenmp -o $@ $^ -lpthread

$(OBJ)/c_api.o: $(REF)/src/c_api.cpp | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

# The reference's own CLI (`niryukti solve/verify`), linked statically against the objects.
$(OUT)/niryukti_ref: $(REF)/app/main.cpp $(OBJS)
	$(CXX) $(CXXFLAGS) -I$(REF)/src $< $(OBJS) -o $@ -lpthread

# Our driver around the reference library: generates the seeded instances in
# memory and calls the reference CPU backend / solve directly.
$(OUT)/ref_driver: oracle/ref_driver.cpp niryukti_amd/csrc/instances.hpp $(OBJS)
	$(CXX) $(CXXFLAGS) -I$(REF)/src -Iniryukti_amd/csrc $< $(OBJS) -o $@ -lpthread

.PHONY: all
