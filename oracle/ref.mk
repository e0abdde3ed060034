# Builds the reference Mandor core.out / setup.out / TFSFdefrag from the C sources where
# they lie under /root/reference (read-only), into oracle/_ref/ only.
# The reference needs MPI: it links against the MPICH that ships in this
# image under /opt/conda (no stand-in headers or libraries are written).
# Used in this container only, to produce and check the golden fixtures.
#
#   make -f oracle/ref.mk            (from the repository root)
REF     ?= /root/reference
SRC     := $(REF)/source
OUT     := oracle/_ref
MPICC   ?= /opt/conda/bin/mpicc
# Flags of the shipped build (Makefile:34-40 of the reference: NaN-filled
# meshes, tracer ids on), with gcc instead of icc and without -ipo.
FLAGS   := -O2 -std=gnu99 -fcommon -I$(SRC) -DLINUX -DMC_FILL_MESH_WITH_NAN \
           -DMC_NO_MPI_ALLOC_MEM -D__need_timespec \
           -DMC_DISABLE_ANSII_ESCAPE_SEQUENCES -DACTIVATE_TRACER=1 -w

CORE_SRC := core/main log test_mesh misc_units misc_cfgReader IO_sys IO_mesh \
  IO_names misc_partition type_reg misc_parameters IO_tecplot IO_fileMap \
  profiler spectr_dump type_mesh2 type_CFile commandLine type_vector core/em \
  core/em_caps parr_meshes misc_socket parr_regLists parr_ghostCells \
  core/em_TFSF core/em_gaussSpot scpic/em_sources scpic/em_mirror type_mesh \
  core/plasma misc_markerPlacer misc_markerPacker core/plasma_rho \
  core/plasma_IO core/plasma_VSP core/plasma_parallel core/plasma_gaussTest \
  core/diag_probe core/diag_WDensity

SETUP_SRC := setup/main misc_units misc_cfgReader type_reg misc_parameters \
  misc_markerPlacer misc_markerPacker log commandLine type_vector IO_sys \
  IO_mesh misc_partition IO_tecplot IO_names IO_fileMap setup/tag \
  setup/tag_units setup/tag_mesh setup/tag_boundary setup/tag_TFSF \
  core/em_TFSF setup/tag_TFSFFaces setup/tag_laser scpic/em_sources \
  scpic/em_mirror setup/tag_plasma setup/tag_DF_uniform setup/tag_point \
  type_mesh setup/tag_EMWave setup/tag_ringDF setup/tag_vShift \
  setup/tag_photoelectrons setup/setup_denavit setup/setup_distrMapper \
  setup/tag_jBalancer setup/tag_twoStream setup/tag_gaussSpot \
  setup/tag_maxwell setup/tag_planePulse setup/tag_plasmaWave \
  setup/tag_EMResonator setup/tag_scissors setup/tag_meanVNoise \
  setup/tag_seedWeibel setup/tag_seed_PITS setup/tag_cluster setup/tag_foil \
  setup/tag_scales setup/tag_trianglePrizm setup/plasma core/plasma_rho \
  profiler parr_meshes parr_regLists misc_socket core/plasma_VSP \
  setup/tag_gradient

DEFRAG_SRC := TFSFdefrag/main log misc_cfgReader type_reg

all: $(OUT)/core.out $(OUT)/setup.out $(OUT)/core_gauss.out $(OUT)/TFSFdefrag.out

$(OUT)/obj/%.o: $(SRC)/%.c
	@mkdir -p $(dir $@)
	$(MPICC) -c $(FLAGS) $< -o $@

$(OUT)/objg/%.o: $(SRC)/%.c
	@mkdir -p $(dir $@)
	$(MPICC) -c $(FLAGS) -DMC_GAUSS_TEST $< -o $@

$(OUT)/core.out: $(addprefix $(OUT)/obj/,$(addsuffix .o,$(CORE_SRC)))
	$(MPICC) $^ -o $@ -lm

$(OUT)/core_gauss.out: $(addprefix $(OUT)/objg/,$(addsuffix .o,$(CORE_SRC)))
	$(MPICC) $^ -o $@ -lm

$(OUT)/setup.out: $(addprefix $(OUT)/obj/,$(addsuffix .o,$(SETUP_SRC)))
	$(MPICC) $^ -o $@ -lm

$(OUT)/TFSFdefrag.out: $(addprefix $(OUT)/obj/,$(addsuffix .o,$(DEFRAG_SRC)))
	$(MPICC) $^ -o $@ -lm

.PHONY: all
