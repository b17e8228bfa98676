# Builds the reference idx2App straight from its own sources under /root/reference
# (no CMake, no generated code). Outputs go only to oracle/_ref/.
# Flags follow the reference's own build: -O3 -mavx2 -fopenmp -Didx2_Avx2.
REF ?= /root/reference/Source
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX ?= g++
CXXFLAGS = -std=c++17 -O3 -mavx2 -fopenmp -Didx2_Avx2 -Wno-format-zero-length -w -fPIC
CORE = Args Assert DataSet FileSystem Format idx2Common idx2Decode idx2ParallelDecode \
       idx2Encode idx2Lookup idx2Read idx2ParallelRead idx2Write idx2SparseBricks \
       InputOutput Logger Memory MemoryMap StackTrace String Utilities VarInt Volume \
       Wavelet Zfp
OBJS = $(addprefix $(OUT)/obj/,$(addsuffix .o,$(CORE))) $(OUT)/obj/idx2.o

all: $(OUT)/idx2App

$(OUT)/obj/%.o: $(REF)/Core/%.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/idx2.o: $(REF)/idx2.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/idx2App: $(OBJS) $(REF)/Applications/idx2App.cpp
	$(CXX) $(CXXFLAGS) $(REF)/Applications/idx2App.cpp $(OBJS) -o $@ -lpthread

clean:
	rm -rf $(OUT)

.PHONY: all clean

# The reference vendors zstd 1.4.5 (Core/zstd/zstd.c).  The oracle linked
# against it reproduces the reference's zstd-compressed sections byte for byte.
$(OUT)/zstd145.o: $(REF)/Core/zstd/zstd.c
	@mkdir -p $(OUT)
	$(CC) -O2 -fPIC -w -c $< -o $@

$(OUT)/libidx2oracle_z145.so: $(OUT)/zstd145.o $(dir $(lastword $(MAKEFILE_LIST)))idx2_oracle.c
	gcc -O2 -g -std=c11 -fPIC -ffp-contract=off -fno-fast-math -fwrapv -w -shared \
	  $(dir $(lastword $(MAKEFILE_LIST)))idx2_oracle.c $(OUT)/zstd145.o -o $@ -lm

all: $(OUT)/libidx2oracle_z145.so
