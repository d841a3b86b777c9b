"""The HIP library on the host: it loads, exports every C-ABI symbol of include/ice3.h,
rejects bad arguments, and its constant/table builder reproduces the compiled
reference's INI_CST + INI_RAIN_ICE (no GPU needed for any of this)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ice3_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def test_header_declares_what_the_library_exports(lib):
    hdr = open(os.path.join(ROOT, "include", "ice3.h")).read()
    declared = set(re.findall(r"\b(ice3_\w+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    hdr = open(os.path.join(ROOT, "include", "ice3.h")).read()
    for cname, py in (("ice3_ice_adjust_in", _lib.AdjustIn), ("ice3_ice_adjust_out", _lib.AdjustOut),
                      ("ice3_rain_ice_in", _lib.RainIn), ("ice3_rain_ice_inout", _lib.RainIO),
                      ("ice3_rain_ice_out", _lib.RainOut), ("ice3_pipeline_in", _lib.PipeIn),
                      ("ice3_pipeline_out", _lib.PipeOut), ("ice3_stats", _lib.Stats),
                      ("ice3_ocnd2_in", _lib.Ocnd2In), ("ice3_ocnd2_out", _lib.Ocnd2Out),
                      ("ice3_rain_ocnd2_in", _lib.RainOcnd2In), ("ice3_hail", _lib.Hail),
                      ("ice3_pipeline_hail", _lib.PipeHail)):
        body = re.search(r"typedef struct \{([^}]*)\} %s;" % cname, hdr).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = re.findall(r"\*\s*(\w+)", body)
        assert tuple(names) == tuple(f for f, _ in py._fields_), cname


def test_bad_arguments_are_rejected_without_a_device(lib):
    p = ctypes.c_void_p
    assert lib.ice3_ice_adjust(None, 0, 4, 90, 4, 50.0, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_rain_ice(None, 0, 4, 90, 4, 50.0, None, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_adjust_rain_ice(None, 0, 4, 90, 4, 50.0, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_ice_adjust_ocnd2(None, 4, 90, 4, 50.0, None, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_ice_adjust_hail(None, 4, 90, 4, 50.0, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_rain_ice_hail(None, 4, 90, 4, 50.0, None, None, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_adjust_rain_ice_hail(None, 4, 90, 4, 50.0, None, None, None, None, None) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_create(None) == _lib.ICE3_EINVAL
    assert lib.ice3_ctx_set_sedimentation(None, 1, 0.8) == _lib.ICE3_EINVAL
    # input validation: argument checks come before any launch
    fp = (p * 2)(None, None)
    rows = (ctypes.c_int * 2)(90, 1)
    cnt = p(None)
    assert lib.ice3_count_nonfinite(0, None, None, 4, 4, None, None) == _lib.ICE3_OK
    assert lib.ice3_count_nonfinite(2, fp, rows, 0, 0, cnt, None) == _lib.ICE3_OK
    assert lib.ice3_count_nonfinite(33, fp, rows, 4, 4, cnt, None) == _lib.ICE3_EINVAL
    assert lib.ice3_count_nonfinite(-1, fp, rows, 4, 4, cnt, None) == _lib.ICE3_EINVAL
    assert lib.ice3_count_nonfinite(2, fp, rows, 4, 3, cnt, None) == _lib.ICE3_EINVAL
    assert lib.ice3_count_nonfinite(2, None, rows, 4, 4, cnt, None) == _lib.ICE3_EINVAL
    assert lib.ice3_count_nonfinite(2, fp, rows, 4, 4, cnt, None) == _lib.ICE3_EINVAL  # NULL counts
    assert lib.ice3_count_nonfinite(2, fp, rows, 4, 4, p(8), None) == _lib.ICE3_EINVAL  # NULL fields
    zero = (ctypes.c_int * 2)(0, 0)
    assert lib.ice3_count_nonfinite(2, fp, zero, 4, 4, p(8), None) == _lib.ICE3_OK  # nothing to read
    neg = (ctypes.c_int * 2)(-1, 0)
    assert lib.ice3_count_nonfinite(2, fp, neg, 4, 4, p(8), None) == _lib.ICE3_EINVAL
    assert b"invalid" in lib.ice3_strerror(_lib.ICE3_EINVAL)
    assert b"sub-step" in lib.ice3_strerror(_lib.ICE3_ESUBSTEP)


def _golden_constants(golden_dir):
    out = {}
    for line in open(os.path.join(golden_dir, "constants.txt")):
        p = line.split()
        out[p[0]] = np.array([float(x) for x in p[2:]])
    return out


def test_constants_match_reference_ini_rain_ice(golden_dir):
    ref = _golden_constants(golden_dir)
    mine = {k: np.array(v) for k, v in _lib.host_constants().items()}
    missing = set(mine) - set(ref)
    assert missing <= {"XT0CRIAUTI"} or not missing
    worst = {}
    for k, v in mine.items():
        if k not in ref:
            continue
        r = ref[k]
        assert v.shape == r.shape, k
        err = np.max(np.abs(v - r) / np.maximum(np.abs(r), 1e-300))
        worst[k] = err
    bad = {k: e for k, e in worst.items() if e > 4e-16 and not k.startswith("XGAMINC")}
    assert not bad, bad
    # GAMMA_INC tables: same algorithm, same abscissae (powers by binary powering)
    for k in ("XGAMINC_RIM1", "XGAMINC_RIM2", "XGAMINC_RIM4"):
        assert worst[k] <= 1e-14, (k, worst[k])


def test_collection_kernel_tables_are_bit_exact(golden_dir):
    ref = _golden_constants(golden_dir)
    mine = _lib.host_constants()
    for k in ("XKER_RACCSS", "XKER_RACCS", "XKER_SACCRG", "XKER_SDRYG", "XKER_RDRYG", "XKER_SWETH",
              "XKER_GWETH", "XKER_RWETH"):
        assert np.array_equal(np.array(mine[k]), ref[k]), k


def test_configuration_objects():
    """The drop-ins accept what the reference's callers pass for the AROME configuration
    (None, "AROME", a Phyex("AROME")-like object with .program, a constants dict for
    RainIce) and reject any other configuration before touching a device."""
    from ice3_amd.ice_adjust import IceAdjust, check_configuration
    from ice3_amd.rain_ice import RainIce

    class Cfg:
        def __init__(self, program):
            self.program = program

    for ok in (None, "AROME", "arome", Cfg("AROME")):
        check_configuration(ok)
    for bad in ("MESO-NH", Cfg("MESO-NH"), 3):
        with pytest.raises(ValueError):
            check_configuration(bad)
    with pytest.raises(ValueError):
        IceAdjust(phyex=Cfg("MESO-NH"))
    with pytest.raises(ValueError):
        RainIce(constants="MESO-NH")

    # a Phyex("AROME")-like object with its switch groups: the AROME values pass (including
    # the Python defaults LSEDIC = LCRIAUTI = 1, which the kernels do not read); a switch
    # changed away from the pinned configuration raises before any device work
    class Group:
        def __init__(self, **kw):
            self.__dict__.update(kw)

    def arome():
        c = Cfg("AROME")
        c.nebn = Group(LSUBG_COND=1, LSIGMAS=1, LSTATNW=0, LHGT_QS=0, CONDENS=0, LAMBDA3=0,
                       FRAC_ICE_ADJUST=0, VSIGQSAT=0.02)
        c.param_icen = Group(LCOND2=0, SUBG_MF_PDF=1, SEDIM=1, LRED=0, MRSTEP=0.0, TSTEP_TS=0.0,
                             LWARM=1, LFEEDBACKT=1, LEVLIMIT=1, LNULLWETG=1, LWETGPOST=1,
                             LCRFLIMIT=1, LSNOW_T=0, LSEDIM_AFTER=0, LDEPOSC=0, SNOW_RIMING=0,
                             SUBG_RC_RR_ACCR=0, SUBG_RR_EVAP=0, SUBG_PR_PDF=0, SUBG_AUCV_RC=1,
                             SUBG_AUCV_RI=0, PRISTINE_ICE="PLAT", LSEDIC=1, LCRIAUTI=1)
        return c

    check_configuration(arome())
    # SEDIM selects the sedimentation scheme of rain_ice: 1 (STAT, AROME) or 0 (SPLI)
    from ice3_amd.ice_adjust import sedimentation_of
    assert sedimentation_of(arome()) == ("STAT", 0.8)
    c = arome()
    c.param_icen.SEDIM, c.param_icen.SPLIT_MAXCFL = 0, 0.5
    check_configuration(c)
    assert sedimentation_of(c) == ("SPLI", 0.5)
    for group, name, value in (("nebn", "LSUBG_COND", 0), ("nebn", "FRAC_ICE_ADJUST", 3),
                               ("param_icen", "SEDIM", 2), ("param_icen", "LCOND2", 1),
                               ("param_icen", "PRISTINE_ICE", "COLU")):
        c = arome()
        setattr(getattr(c, group), name, value)
        with pytest.raises(ValueError, match=name):
            check_configuration(c)
        with pytest.raises(ValueError, match=name):
            IceAdjust(phyex=c)
