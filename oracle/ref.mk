# Builds the reference's own CPU operators (reference/db_operators/*.cpp,
# compiled where they lie under the reference tree) into oracle/_ref/ as
# ctypes-loadable libraries, through oracle/ref_harness.cpp, and the PIM
# library's hash mixers (pimdal/hash/hash_func.h, plain C) through
# oracle/ref_hash_harness.c.  Test infrastructure: these pin the CPU
# restatement in oracle.cpp.
#
# Needs: the reference tree, pyarrow's bundled Arrow C++ (headers and
# libarrow.so.2500) and five source files of the abseil copy vendored in the
# reference (flat_hash_set/flat_hash_map).  C++20 because Arrow 25's headers
# need std::span (the reference pins C++17 against Arrow 12).
#
#   make -C oracle -f ref.mk
REF ?= /root/reference/reference/db_operators
PIMHASH ?= /root/reference/pimdal/hash
PYARROW ?= $(shell python3 -c "import pyarrow, os; print(os.path.dirname(pyarrow.__file__))")
AB := $(REF)/abseil-cpp
OUT := _ref
N := 800000
CXX ?= g++
FLAGS := -std=c++20 -O2 -fopenmp -fPIC -shared -w -DREF_N=$(N) -DBUFFER_SIZE=$(N) \
         -I$(REF) -I$(AB) -I$(PYARROW)/include
LIBS := -L$(PYARROW) -l:libarrow.so.2500 -Wl,-rpath,$(PYARROW) -Wl,--no-undefined
ABSL := $(AB)/absl/container/internal/raw_hash_set.cc $(AB)/absl/hash/internal/hash.cc \
        $(AB)/absl/hash/internal/city.cc $(AB)/absl/hash/internal/low_level_hash.cc \
        $(AB)/absl/base/internal/raw_logging.cc

all: $(OUT)/ref_hash_func.so $(OUT)/ref_select.so $(OUT)/ref_sort.so $(OUT)/ref_hash_join.so $(OUT)/ref_aggregate.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/ref_select.so: ref_harness.cpp | $(OUT)
	$(CXX) $(FLAGS) -DREF_SELECT -DREF_SRC='"select.cpp"' ref_harness.cpp -o $@ $(LIBS)

$(OUT)/ref_sort.so: ref_harness.cpp | $(OUT)
	$(CXX) $(FLAGS) -DREF_SORT -DREF_SRC='"sort.cpp"' ref_harness.cpp -o $@ $(LIBS)

$(OUT)/ref_hash_join.so: ref_harness.cpp | $(OUT)
	$(CXX) $(FLAGS) -DREF_HASH_JOIN -DREF_SRC='"hash_join.cpp"' ref_harness.cpp $(ABSL) -o $@ $(LIBS)

$(OUT)/ref_aggregate.so: ref_harness.cpp | $(OUT)
	$(CXX) $(FLAGS) -DREF_AGGREGATE -DREF_SRC='"aggregate.cpp"' ref_harness.cpp $(ABSL) -o $@ $(LIBS)

$(OUT)/ref_hash_func.so: ref_hash_harness.c | $(OUT)
	$(CC) -O2 -fPIC -shared -w -I$(PIMHASH) ref_hash_harness.c -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
