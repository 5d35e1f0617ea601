# Builds the reference Namics CPU program from its own sources, where they lie
# under /root/reference, into oracle/_ref/ (git-ignored).  Used only as a
# checker: to generate golden fixtures in tests/golden/ and to time the
# reference CPU path.  Plain -O2 (not the reference Makefile's -Ofast), so
# floating point follows IEEE semantics.
REF      ?= /root/reference
OUT      := $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
CXXFLAGS := -O2 -std=c++14 -w -I$(REF)/src
SRCS     := $(shell find $(REF)/src -type f -name '*.cpp' -not -path '*/Eigen/*')
OBJS     := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/namics

$(OUT)/namics: $(OBJS)
	$(CXX) -o $@ $^ -lm -lpthread

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c -o $@ $<

.PHONY: all
