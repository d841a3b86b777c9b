# Builds the PHYEX CY50T1 Fortran reference in FP64 (-fdefault-real-8) as a shared
# library, straight from the sources where they lie under $(REF), plus the
# build-owned driver oracle/ref_driver.F90.  Outputs go to oracle/_ref/ only.
#
# The source list and its order follow the reference's own CMakeLists.txt:62-304
# (stubs, modd_*, modi_*, mode_*, functions, main routines), minus the C_FLOAT
# bridge (bridge/phyex_bridge.F90 does not compile in FP64) and the shallow
# convection scheme, which is not on the ice_adjust/rain_ice path.
#
# Two variants from the same sources:
# * VARIANT=parity (default, oracle/_ref/): -O2 for the x86-64 baseline, which has no
#   FMA, so flang cannot contract a*b+c and the reference arithmetic stays the plain IEEE
#   sequence.  It writes the golden vectors and pins the C restatement.
# * VARIANT=fast (oracle/_ref/fast/): -O3 for x86-64-v3 (AVX2 + FMA), timing only: the
#   optimised CPU baseline.  The library is built on one machine and timed on the host of
#   the machine with the GPU, so it targets an instruction set both hosts have rather than
#   -march=native of the build machine.
#
# Usage (where /root/reference holds the sources; the timing machine receives the built
# library, not the sources):
#   make -f oracle/ref_build.mk [VARIANT=fast]

REF      ?= /root/reference/PHYEX-IAL_CY50T1
VARIANT  ?= parity
ifeq ($(VARIANT),fast)
OUT      := oracle/_ref/fast
OPT      := -O3 -march=x86-64-v3
else
OUT      := oracle/_ref
OPT      := -O2
endif
OBJ      := $(OUT)/obj
MOD      := $(OUT)/mod
FC       := flang
FFLAGS   := $(OPT) -fPIC -cpp -fdefault-real-8 -fdefault-double-8 -module-dir $(MOD) -I$(MOD) -I$(REF)/stubs/include -I$(REF)/micro

STUB := stubs/parkind1.F90 stubs/ec_parkind.F90 stubs/yomhook.F90 stubs/yomlun.F90 stubs/abor1.F90 \
        stubs/fmlook_ll.F90 stubs/posnam.F90
AUX_MODD := aux/modd_precision.F90 aux/modd_parameters.F90 aux/modd_cst.F90 aux/modd_dimphyexn.F90 \
        aux/modd_field.F90 aux/modd_io.F90 aux/modd_argslist_ll.F90 aux/modd_budget.F90 aux/modd_conf.F90 \
        aux/modd_les.F90 aux/modd_misc.F90 aux/modd_nsv.F90
MICRO_MODD := micro/modd_tiwmx.F90 micro/modd_rain_ice_descrn.F90 micro/modd_rain_ice_paramn.F90 \
        micro/modd_param_icen.F90 micro/modd_nebn.F90 micro/modd_cloudparn.F90 micro/modd_fields_address.F90
TURB_MODD := turb/modd_cturb.F90 turb/modd_turbn.F90 turb/modd_param_mfshalln.F90
CONV_MODD := conv/modd_convpar.F90 conv/modd_convpar_shal.F90 conv/modd_convparext.F90
PHYEX := aux/modd_phyex.F90
MODI := aux/modi_gamma.F90 aux/modi_gamma_inc.F90 aux/modi_general_gamma.F90 aux/modi_second_mnh.F90 \
        aux/modi_shuman.F90 aux/modi_gradient_m.F90 aux/modi_gradient_u.F90 aux/modi_gradient_v.F90 \
        aux/modi_gradient_w.F90 aux/modi_ini_phyex.F90 micro/modi_ice_adjust.F90 micro/modi_condensation.F90 \
        micro/modi_rain_ice.F90
AUX_MODE := aux/mode_msg.F90 aux/mode_mppdb.F90 aux/mode_ll.F90 aux/mode_thermo.F90 aux/mode_check_nam_val.F90 \
        aux/mode_fill_dimphyexn.F90 aux/mode_posnam_phy.F90 aux/mode_shuman_phy.F90 aux/mode_sources_neg_correct.F90 \
        aux/mode_ini_cst.F90 aux/mode_io_field_write.F90 aux/mode_io_field_write_phy.F90 aux/mode_argslist_ll_phy.F90 \
        aux/mode_gradient_m_phy.F90 aux/mode_gradient_u_phy.F90 aux/mode_gradient_v_phy.F90 \
        aux/mode_gradient_w_phy.F90 aux/mode_mnh_zwork.F90
AUX_FUNC := aux/shuman.F90 aux/gamma.F90 aux/gamma_inc.F90 aux/general_gamma.F90 aux/second_mnh.F90 \
        aux/gradient_m.F90 aux/gradient_u.F90 aux/gradient_v.F90 aux/gradient_w.F90 aux/ini_phyex.F90 \
        aux/tools.F90 aux/bitrep.F90
MICRO_MODE := micro/mode_tiwmx.F90 micro/mode_tiwmx_fun.F90 micro/mode_tiwmx_tab.F90 micro/mode_qsatmx_tab.F90 \
        micro/mode_icecloud.F90 micro/mode_ini_tiwmx.F90 micro/mode_rrcolss.F90 micro/mode_rzcolx.F90 \
        micro/mode_rscolrg.F90 micro/mode_read_xker_raccs.F90 micro/mode_read_xker_sdryg.F90 \
        micro/mode_read_xker_rdryg.F90 micro/mode_read_xker_sweth.F90 micro/mode_read_xker_gweth.F90 \
        micro/mode_read_xker_rweth.F90 micro/mode_ini_rain_ice.F90 micro/mode_ice4_rainfr_vert.F90 \
        micro/mode_ice4_compute_pdf.F90 micro/mode_ice4_sedimentation_stat.F90 micro/mode_ice4_pack.F90 \
        micro/mode_ice4_correct_negativities.F90 micro/mode_ice4_budgets.F90 micro/mode_ice4_fast_rg.F90 \
        micro/mode_ice4_fast_rh.F90 micro/mode_ice4_fast_ri.F90 micro/mode_ice4_fast_ri_rs.F90 \
        micro/mode_ice4_fast_rs.F90 micro/mode_ice4_rimltc.F90 micro/mode_ice4_rrhong.F90 micro/mode_ice4_slow.F90 \
        micro/mode_ice4_stepping.F90 micro/mode_ice4_tendencies.F90 micro/mode_ice4_warm.F90 \
        turb/mode_ini_mfshall.F90 turb/mode_ini_turb.F90
MICRO_FUNC := micro/hypser.f90 micro/hypgeo.F90 micro/momg.F90
MAIN := micro/condensation.F90 micro/ice_adjust.F90 micro/rain_ice.F90

SRCS_CMAKE := $(STUB) $(AUX_MODD) $(MICRO_MODD) $(TURB_MODD) $(CONV_MODD) $(PHYEX) $(MODI) $(AUX_MODE) \
        $(AUX_FUNC) $(MICRO_MODE) $(MICRO_FUNC) $(MAIN)
# module-dependency order (CMake's scanner does this for the reference's own build)
SRCS := $(shell python3 oracle/ref_order.py $(REF) $(SRCS_CMAKE))
OBJS := $(addprefix $(OBJ)/,$(addsuffix .o,$(basename $(notdir $(SRCS)))))

.NOTPARALLEL:

all: $(OUT)/libphyex_ref.so

objs: $(OBJS)

$(OBJ) $(MOD):
	mkdir -p $@

# Objects are built one at a time in SRCS order (.NOTPARALLEL), so every module
# file exists before the first source that uses it.
define compile_rule
$(OBJ)/$(basename $(notdir $(1))).o: $(REF)/$(1) | $(OBJ) $(MOD)
	$(FC) $(FFLAGS) -c $$< -o $$@
endef
$(foreach s,$(SRCS),$(eval $(call compile_rule,$(s))))

$(OBJ)/ref_driver.o: oracle/ref_driver.F90 $(OBJS) | $(OBJ) $(MOD)
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/libphyex_ref.so: $(OBJS) $(OBJ)/ref_driver.o
	$(FC) -shared -fPIC -o $@ $^ -Wl,-rpath,$(shell dirname $(shell which $(FC)))/../lib

clean:
	rm -rf $(OUT)
