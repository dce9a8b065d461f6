# Builds the reference x2dhf program (Fortran 90) straight from its sources
# under /root/reference/src with AMD flang, without the reference's CMake.
# Output goes to oracle/_ref/ only (git-ignored); it is the checker used to
# pin the restated oracle and to produce golden fixtures, never the product.
#
#   make -f oracle/ref.mk            # serial build   -> oracle/_ref/x2dhf
#   make -f oracle/ref.mk omp        # OpenMP build   -> oracle/_ref/x2dhf_omp
#   make -f oracle/ref.mk pthread    # pthreads build -> oracle/_ref/x2dhf_pt
#     (the reference's USE_PTHREAD variant: Fortran with -DPTHREAD, the MCSOR
#      and potential threads of src/c with gcc -pthread)
#   make -f oracle/ref.mk tpool      # pthreads thread-pool build -> oracle/_ref/x2dhf_tp
#     (USE_TPOOL: the same relaxation routines on persistent pthreads)
# The USE_PTHREAD build aborts ("stack smashing detected") as soon as it
# relaxes with more than one MCSOR thread: its pthread_join writes a void *
# into an int (src/c/mcsor_pthread.c:83-87, src/c/coulExch_pthread.c:119-120).
# The thread-pool build joins with NULL and runs, but its orbital sweeps
# diverge on the served-stream decks (orbital norms of 1e14 and then NaN in
# the iteration after its MCSOR threads start), so the CPU baseline of the
# stream is the OpenMP build on the same number of threads.
#
# Internal BLAS is used (the reference's USE_BLAS=OFF), so every operation on
# the relaxation path is the reference's own code.

REFSRC  ?= /root/reference/src
OUT     ?= oracle/_ref
REF_FC  ?= /opt/rocm/lib/llvm/bin/flang
FFLAGS  ?= -O2
# module dependency order of the reference's src/CMakeLists.txt source list
SRCS = dgamit.F params.F90 blas.F90 card.f90 discret.f90 checkOrbSym.f90 memory.f90 \
  solver.f90 commons.F90 utils.f90 sapPotential.f90 sharedMemory.F90 elocc.f90 \
  coulombExchangePotentials.f90 data4II.f90 dateTime.f90 inout.F90 nabla.f90 dftvxc.f90 \
  dointerp.f90 exchContribs.f90 orbAsympt.f90 printUtils.f90 scfshr.f90 sort.f90 detect.F90 \
  dftexc.F90 dumpDataToDisk.F90 diskInterface.F90 diskInterfaceMisc.F90 integrals.F90 \
  lagrangeMultipliers.F90 totalEnergy.F90 totalEnergyLXC.F90 sormcsor.F90 fock.F90 \
  relaxOrbs.F90 scfUtils.F90 normOrtho.F90 relaxPots.F90 doSCF.F90 initCBAllocArrays.F90 \
  prepGauss.F90 storedOrbitals.F90 initOrbitalsPotentials.F90 initVariables.F90 \
  inputInterface.F90 machine.F90 printCase.F90 summary.F90 x2dhf.F90

.PHONY: serial omp pthread tpool all
serial: $(OUT)/x2dhf
omp: $(OUT)/x2dhf_omp
pthread: $(OUT)/x2dhf_pt
tpool: $(OUT)/x2dhf_tp
all: serial omp pthread tpool

CSRCS = sorpt.c mcsor_pthread.c coulExch_pthread.c relax_single_pot_pthread.c mcsor_single_colour_pthread.c
TSRCS = sorpt.c mcsor_tpool.c coulExch_tpool.c relax_single_pot_tpool.c mcsor_single_colour_tpool.c
REF_CC ?= gcc
CFLAGS ?= -O2

# Sources are compiled one at a time in dependency order (module files are
# produced as a side effect), so a single recipe per variant is enough.
$(OUT)/x2dhf: $(addprefix $(REFSRC)/,$(SRCS))
	mkdir -p $(OUT)/serial
	cd $(OUT)/serial && for f in $(SRCS); do \
	  $(REF_FC) $(FFLAGS) -I$(REFSRC) -c $(REFSRC)/$$f || exit 1; done
	cd $(OUT)/serial && $(REF_FC) $(FFLAGS) -o ../x2dhf *.o

$(OUT)/x2dhf_omp: $(addprefix $(REFSRC)/,$(SRCS))
	mkdir -p $(OUT)/omp
	cd $(OUT)/omp && for f in $(SRCS); do \
	  $(REF_FC) $(FFLAGS) -fopenmp -I$(REFSRC) -c $(REFSRC)/$$f || exit 1; done
	cd $(OUT)/omp && $(REF_FC) $(FFLAGS) -fopenmp -o ../x2dhf_omp *.o

$(OUT)/x2dhf_pt: $(addprefix $(REFSRC)/,$(SRCS)) $(addprefix $(REFSRC)/c/,$(CSRCS))
	mkdir -p $(OUT)/pt
	cd $(OUT)/pt && for f in $(SRCS); do \
	  $(REF_FC) $(FFLAGS) -DPTHREAD -I$(REFSRC) -c $(REFSRC)/$$f || exit 1; done
	cd $(OUT)/pt && for f in $(CSRCS); do \
	  $(REF_CC) $(CFLAGS) -pthread -DPTHREAD -I$(REFSRC) -I$(REFSRC)/c -c $(REFSRC)/c/$$f || exit 1; done
	cd $(OUT)/pt && $(REF_FC) $(FFLAGS) -o ../x2dhf_pt *.o -lpthread

$(OUT)/x2dhf_tp: $(addprefix $(REFSRC)/,$(SRCS)) $(addprefix $(REFSRC)/c/,$(TSRCS))
	mkdir -p $(OUT)/tp
	cd $(OUT)/tp && for f in $(SRCS); do \
	  $(REF_FC) $(FFLAGS) -DTPOOL -I$(REFSRC) -c $(REFSRC)/$$f || exit 1; done
	cd $(OUT)/tp && for f in $(TSRCS); do \
	  $(REF_CC) $(CFLAGS) -pthread -DTPOOL -I$(REFSRC) -I$(REFSRC)/c -c $(REFSRC)/c/$$f || exit 1; done
	cd $(OUT)/tp && $(REF_FC) $(FFLAGS) -o ../x2dhf_tp *.o -lpthread

# Golden-vector harness: oracle/ref_harness.f90 linked against the OpenMP
# objects of the reference (everything except its main program).
.PHONY: harness
harness: $(OUT)/ref_harness
$(OUT)/ref_harness: oracle/ref_harness.f90 $(OUT)/x2dhf_omp
	cd $(OUT)/omp && $(REF_FC) $(FFLAGS) -fopenmp -I. -c ../../ref_harness.f90 -o ../ref_harness.o
	$(REF_FC) $(FFLAGS) -fopenmp -o $@ $(OUT)/ref_harness.o $(filter-out $(OUT)/omp/x2dhf.o,$(wildcard $(OUT)/omp/*.o))
