# Builds oracle/_ref/libmdtkref.so: the reference's own GSL-free translation
# units, compiled in place from the reference tree, plus ref_harness.cpp (C
# entry points for the CPU tests).  Test infrastructure only; nothing is
# copied from the reference tree and nothing here is part of the product.
#   make -f oracle/ref.mk            (REF defaults to /root/reference)
# The spline-based potentials need GSL (mdtk/SplineAux.cxx:34), which is not
# installed, so REBO/AIREBO/ETors/Brenner/FLJ/Ackland/TightBinding and
# SimLoop are not part of this build.
REF ?= /root/reference
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref
CXX ?= g++
CXXFLAGS ?= -O2 -std=gnu++98 -fPIC -ffp-contract=off -fno-fast-math -w
# the harness is C++17; the reference units keep their own C++98 dialect
HFLAGS ?= -O2 -std=c++17 -fPIC -ffp-contract=off -fno-fast-math -Wall

REF_SRC := yaatk/yaatk.cxx yaatk/Exception.cxx yaatk/config.cxx \
  mdtk/Atom.cxx mdtk/AtomsContainer.cxx mdtk/Vector3D.cxx mdtk/consts.cxx mdtk/config.cxx \
  mdtk/Exception.cxx mdtk/tools.cxx mdtk/release_info.cxx \
  mdtk/potentials/AtomsPair.cxx mdtk/potentials/FGeneral.cxx mdtk/potentials/NeighbourList.cxx \
  mdtk/potentials/FProxy.cxx mdtk/potentials/pairwise/FPairwise.cxx \
  mdtk/potentials/pairwise/FBZL.cxx mdtk/potentials/pairwise/FBM.cxx
REF_OBJ := $(addprefix $(OUT)/,$(subst /,_,$(REF_SRC:.cxx=.o)))

$(OUT)/libmdtkref.so: $(REF_OBJ) $(OUT)/ref_harness.o
	$(CXX) -shared -o $@ $^

$(OUT)/ref_harness.o: $(HERE)ref_harness.cpp | $(OUT)
	$(CXX) $(HFLAGS) -isystem $(REF) -c $< -o $@

define refobj
$(OUT)/$(subst /,_,$(1:.cxx=.o)): $(REF)/$(1) | $(OUT)
	$$(CXX) $$(CXXFLAGS) -I$(REF) -c $$< -o $$@
endef
$(foreach s,$(REF_SRC),$(eval $(call refobj,$(s))))

$(OUT):
	mkdir -p $@

clean:
	rm -rf $(OUT)
