# Test infrastructure: compiles the reference SRW sources where they lie
# (/root/reference/cpp/src) into oracle/_ref/ and links the command-line runner
# oracle/ref_run.cpp against them.  Nothing is copied out of the reference and
# nothing here is part of the library.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags follow the reference's own gcc build (cpp/gcc/Makefile: -O3 -fPIC and
# its SRW_SRC_DEF defines, MODE=0).  FFTW is not installed in this image (the
# reference ships it only as a source tarball with a configure step), so the
# runner is linked with its FFTW and f2c-I/O symbols left unresolved and bound
# lazily; the runner only calls entry points that never reach them.
REF   ?= /root/reference/cpp/src
OUT   := oracle/_ref
OBJD  := $(OUT)/obj
CXX   ?= g++
CC    ?= gcc
DEFS  := -D_GNU_SOURCE -D__USE_XOPEN2K8 -DFFTW_ENABLE_FLOAT -D_GM_WITHOUT_BASE -DSRWLIB_STATIC -DNO_TIMER \
         -DANSI_DECLARATORS -DTRILIBRARY -DLINUX -D_FFTW3
INCS  := -I$(REF)/core -I$(REF)/lib -I$(REF)/ext/auxparse -I$(REF)/ext/genmath
FLAGS := -O3 -fPIC -w

CORE  := srclcuti srcradint srctrjdt sremitpr srgsnbm srgtrjdt srisosrc srmagcnt srmagfld srmatsta sroptapt \
         sroptcnt sroptdrf sroptel2 sroptel3 sroptelm sroptfoc sroptgrat sroptgtr sropthck sroptcryst sroptmat \
         sroptpsh sroptshp sroptsmr sroptwgr sroptzp sroptzps srpersto srpowden srprdint srprgind srpropme \
         srptrjdt srradinc srradint srradmnp srradstr srremflp srsase srsend srstowig srsysuti srthckbm \
         srthckbm2 srtrjaux srtrjdat srtrjdat3d
GMATH := gmfft gmfit gminterp gmmeth gmtrans
GENES := all_com check diagno esource field incoherent initrun input loadbeam loadrad magfield main math mpi \
         output partsim pushp rpos scan source stepz string tdepend timerec track
F2C   := D_CNJG

OBJS  := $(addprefix $(OBJD)/,$(addsuffix .o,$(CORE) $(GMATH) auxparse srerror srwlib)) \
         $(addprefix $(OBJD)/g_,$(addsuffix .o,$(GENES))) $(addprefix $(OBJD)/f2c_,$(addsuffix .o,$(F2C)))

all: $(OUT)/ref_run

$(OBJD):
	mkdir -p $@

$(OBJD)/%.o: $(REF)/core/%.cpp | $(OBJD)
	$(CXX) $(FLAGS) $(DEFS) $(INCS) -c $< -o $@
$(OBJD)/%.o: $(REF)/ext/genmath/%.cpp | $(OBJD)
	$(CXX) $(FLAGS) $(DEFS) $(INCS) -c $< -o $@
$(OBJD)/%.o: $(REF)/ext/auxparse/%.cpp | $(OBJD)
	$(CXX) $(FLAGS) $(DEFS) $(INCS) -c $< -o $@
$(OBJD)/%.o: $(REF)/lib/%.cpp | $(OBJD)
	$(CXX) $(FLAGS) $(DEFS) $(INCS) -c $< -o $@
$(OBJD)/g_%.o: $(REF)/ext/genesis/genesis_july08/%.c | $(OBJD)
	$(CC) -O2 -fPIC -w -D_GNU_SOURCE -DLINUX -I$(REF)/ext/genmath -c $< -o $@
$(OBJD)/f2c_%.o: $(REF)/ext/f2c_min/%.C | $(OBJD)
	$(CC) -O2 -fPIC -w -x c -I$(REF)/ext/genmath -c $< -o $@

$(OUT)/libsrw_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -Wl,-z,lazy -lm

$(OUT)/ref_run: oracle/ref_run.cpp $(OUT)/libsrw_ref.so
	$(CXX) -O2 $(DEFS) $(INCS) oracle/ref_run.cpp -o $@ -L$(OUT) -lsrw_ref -Wl,-rpath,'$$ORIGIN' \
	    -Wl,--allow-shlib-undefined -Wl,-z,lazy -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
