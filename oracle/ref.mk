# Builds the parts of the reference that compile from their own source files
# alone into oracle/_ref/ (checker only; never shipped, never copied).
#
# simple_physics_v5_1.f90 is a self-contained subroutine (no modules, no
# external libraries), so it compiles as is. The dycore itself (dyn_run.f90 &c.)
# needs SHTns, FFTW3 and LAPACK, which the image lacks: not built here.
REF ?= /root/reference/src
FC := /opt/rocm/bin/amdflang
OUT := _ref
all: $(OUT)/libref_simple_physics.so
$(OUT)/libref_simple_physics.so: $(REF)/simple_physics_v5_1.f90
	mkdir -p $(OUT)/mod
	$(FC) -O2 -fPIC -shared -J $(OUT)/mod -o $@ $<
clean:
	rm -rf $(OUT)
.PHONY: all clean
