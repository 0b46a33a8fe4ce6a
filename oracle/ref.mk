# Builds the reference's own Fortran/C routines for the checker, from the
# reference's sources where they lie, into oracle/_ref/ (never committed):
#   libbrawl_ref.so  the energy, initial-set-up, random-site and
#                    radial-density routines behind C entry points
#                    (oracle/ref_shim.f90), for the oracle's cross-checks
#   libdrop_in.so    host code that rebinds the reference's procedure
#                    pointers to the GPU adaptors (brawl_amd/fortran) and
#                    checks them against the reference's routines
# Modules that need NetCDF or a flang-built MPI module (metropolis.F90,
# nested_sampling.f90, wang-landau.F90, main.F90) are not built; the shim
# restates their sampling loops (Metropolis sweeps, Wang-Landau sweeps and
# window entry, the nested-sampling loop) over the built modules' routines.
REF ?= /root/reference/src
AMDFLANG ?= /opt/rocm/lib/llvm/bin/amdflang
FC := $(AMDFLANG)
CC = gcc
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
OBJ := $(OUT)/obj
FFLAGS = -O2 -cpp -fPIC -ffp-contract=off -module-dir $(OBJ) -I$(OBJ)
CFLAGS = -O2 -fPIC -ffp-contract=off
MODS = kinds constants shared_data derived_types c_functions bw_hamiltonian random_site \
       comms command_line display io analytics initialise
REF_OBJS = $(OBJ)/mt19937ar.o $(addprefix $(OBJ)/,$(addsuffix .o,$(MODS)))
LIBDIR := $(HERE)../brawl_amd

all: $(OUT)/libbrawl_ref.so $(OUT)/libdrop_in.so

$(OBJ)/mt19937ar.o: $(REF)/mt19937ar.c
	@mkdir -p $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

# the reference's modules, in dependency order
$(OBJ)/%.o: $(REF)/%.f90
	@mkdir -p $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/%.o: $(REF)/%.F90
	@mkdir -p $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/ref_shim.o: $(HERE)ref_shim.f90 $(REF_OBJS)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/brawl_hip.o: $(HERE)../brawl_amd/fortran/brawl_hip.f90
	@mkdir -p $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/brawl_hip_adaptors.o: $(HERE)../brawl_amd/fortran/brawl_hip_adaptors.f90 $(OBJ)/brawl_hip.o $(REF_OBJS)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/drop_in.o: $(HERE)../tests/fortran/drop_in.f90 $(OBJ)/brawl_hip_adaptors.o
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/libbrawl_ref.so: $(OBJ)/ref_shim.o
	$(FC) -shared -o $@ $(OBJ)/ref_shim.o $(REF_OBJS)

$(OUT)/libdrop_in.so: $(OBJ)/drop_in.o $(LIBDIR)/libbrawl_hip.so
	$(FC) -shared -o $@ $(OBJ)/drop_in.o $(OBJ)/brawl_hip_adaptors.o $(OBJ)/brawl_hip.o $(REF_OBJS) \
	  -L$(LIBDIR) -lbrawl_hip -Wl,-rpath,'$$ORIGIN/../../brawl_amd'

clean:
	rm -rf $(OUT)

# module dependencies (each module's .mod comes with its .o)
$(OBJ)/constants.o: $(OBJ)/kinds.o
$(OBJ)/shared_data.o: $(OBJ)/constants.o
$(OBJ)/derived_types.o: $(OBJ)/shared_data.o
$(OBJ)/c_functions.o: $(OBJ)/derived_types.o
$(OBJ)/bw_hamiltonian.o: $(OBJ)/c_functions.o
$(OBJ)/random_site.o: $(OBJ)/bw_hamiltonian.o
$(OBJ)/comms.o: $(OBJ)/random_site.o
$(OBJ)/command_line.o: $(OBJ)/comms.o
$(OBJ)/display.o: $(OBJ)/command_line.o
$(OBJ)/io.o: $(OBJ)/display.o
$(OBJ)/analytics.o: $(OBJ)/io.o
$(OBJ)/initialise.o: $(OBJ)/analytics.o

.PHONY: all clean
