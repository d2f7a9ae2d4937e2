"""The C-ABI library loads and exports every entry point include/mbpol_hip.h
declares; host-side argument checking works without a GPU.  CPU only."""
import ctypes
import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'mbpol_hip.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mbpol_[a-z_0-9]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    import mbpol_openmm_plugin_amd as m
    L = m.library()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_system_validation_without_gpu():
    import mbpol_openmm_plugin_amd as m
    L = m.library()
    h = ctypes.c_void_p()
    assert L.mbpol_system_create(ctypes.c_int32(0), ctypes.byref(h)) == 1
    assert L.mbpol_system_create(ctypes.c_int32(6), ctypes.byref(h)) == 0
    bad = (ctypes.c_int32 * 3)(0, 1, 99)
    assert L.mbpol_system_set_twobody(h, 1, bad, 1, ctypes.c_double(0.65)) == 1
    assert b'out of range' in L.mbpol_last_error()
    good = (ctypes.c_int32 * 6)(0, 1, 2, 3, 4, 5)
    assert L.mbpol_system_set_twobody(h, 2, good, 7, ctypes.c_double(0.65)) == 1
    assert L.mbpol_system_set_twobody(h, 2, good, 1, ctypes.c_double(0.65)) == 0
    L.mbpol_system_destroy(h)


def test_python_force_api_mirrors_reference_defaults():
    import mbpol_openmm_plugin_amd as m
    el = m.MBPolElectrostaticsForce()
    # openmmapi/src/MBPolElectrostaticsForce.cpp:43-49
    assert el.getNonbondedMethod() == el.NoCutoff
    assert el.getCutoffDistance() == 0.9
    assert el.getEwaldErrorTolerance() == 1e-4
    assert el.getMutualInducedMaxIterations() == 200
    assert el.getMutualInducedTargetEpsilon() == 1e-7
    assert el.getTholeParameters() == [0.4, 0.4, 0.055, 0.626, 0.055]
    assert el.getPmeBSplineOrder() == 5
    tb = m.MBPolTwoBodyForce()
    assert tb.getCutoff() == 1e10 and tb.getNonbondedMethod() == tb.NoCutoff
    i = tb.addParticle([0, 1, 2])
    assert tb.getParticleParameters(i) == [0, 1, 2] and tb.getNumMolecules() == 1


def test_context_without_device_fails_loudly():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() > 0:
        pytest.skip('a GPU is visible')
    sys_ = m.forcefield.create_system(3, nonbonded_method='CutoffNonPeriodic')
    with pytest.raises(m.MBPolError) as e:
        m.Context(sys_)
    assert e.value.code == 7


def test_group_validation_without_gpu():
    """Independent groups: malformed offsets are refused by the system call, and a
    system that does not have the water layout is refused before any GPU work."""
    import mbpol_openmm_plugin_amd as m
    L = m.library()
    h = ctypes.c_void_p()
    assert L.mbpol_system_create(ctypes.c_int32(8), ctypes.byref(h)) == 0
    for bad in ((1, 2), (0, 0), (0, 2, 1)):
        arr = (ctypes.c_int32 * len(bad))(*bad)
        assert L.mbpol_system_set_groups(h, len(bad) - 1, arr) == 1, bad
    L.mbpol_system_destroy(h)
    # two waters, groups of 1 + 1, but a periodic box: refused (the check runs before the device query)
    sysm = m.forcefield.create_system(2, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                      threebody_cutoff=1e10, nonbonded_cutoff=1e3)
    sysm.setWaterGroups([0, 1, 3])  # three waters claimed, two present
    with pytest.raises(m.MBPolError) as e:
        m.Context(sysm)
    assert 'groups' in str(e.value)
