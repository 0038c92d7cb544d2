# Builds the CPU oracle (test infrastructure): a shared library for the tests
# and a command-line binary for the benchmark's CPU-baseline leg.
#   make -f oracle/oracle.mk -j8
OUT := oracle/lib
CXX := g++
CXXFLAGS := -std=c++20 -O2 -fopenmp -fPIC -Wall -Wextra -Ioracle -Iniryukti_amd/csrc

all: $(OUT)/liboracle_pdhg.so $(OUT)/pdhg_oracle

$(OUT):
	mkdir -p $(OUT)

$(OUT)/liboracle_pdhg.so: oracle/pdhg_oracle.cpp oracle/pdhg_oracle.h | $(OUT)
	$(CXX) $(CXXFLAGS) -shared $< -o $@

$(OUT)/pdhg_oracle: oracle/oracle_main.cpp oracle/pdhg_oracle.cpp oracle/pdhg_oracle.h niryukti_amd/csrc/instances.hpp | $(OUT)
	$(CXX) $(CXXFLAGS) oracle/oracle_main.cpp oracle/pdhg_oracle.cpp -o $@

.PHONY: all
