"""The oracle's strain-rate partner functions against the reference's own
(Node.cpp:1892-2490), bit for bit.

tests/golden/kat_partner.npz holds known answers that the reference build
(oracle/_ref, its unmodified Node.cpp) produced on standalone node clusters:
a centre node, its six axis neighbours, their partners on the other level and
second neighbours, every node with random populations in both arrays and in
its pre-collision copy.  The level-0 flag was true and the level-1 flag false,
so reading the wrong array, dropping the time average or using the wrong
stencil distance changes the result; the planted mistakes below check that.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from make_golden import DX, params_of

L = ol.lib()
D = ol.dptr
K = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat_partner.npz"))
NS = int(K["nsamples"])
SPEC = json.loads(str(K["spec"]))

IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)
KP = ctypes.POINTER(ol.Consts)
L.acpo_strain_coarse_partner.argtypes = [KP, ctypes.c_double, IP, DP, DP, DP, DP]
L.acpo_strain_half_coarse_partner.argtypes = [KP, ctypes.c_double, IP, DP, DP, DP, DP, DP]
L.acpo_strain_fine_partner.argtypes = [KP, ctypes.c_double, DP, DP, DP]
L.acpo_strain_direct_coupling.argtypes = [KP, ctypes.c_int, ctypes.c_double, IP, DP, DP, DP, DP, DP]


def _arr(name, width):
    return np.ascontiguousarray(K[name].reshape(NS, -1, width) if width else K[name].reshape(NS, -1))


NOW, OTHER, PRE = _arr("in_now", 19), _arr("in_other", 19), _arr("in_pre", 19)
FINE, FINE_OTHER = _arr("in_fine", 19), _arr("in_fine_other", 19)
MASK = K["in_mask"].reshape(NS, 6).astype(np.int32)


def _consts():
    return ol.consts(params_of(SPEC))


def _s6_from_ref(t9):
    t = t9.reshape(3, 3)
    assert t[1, 0] == 0 and t[2, 0] == 0 and t[2, 1] == 0  # the reference fills the upper triangle
    return np.array([t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[0, 2], t[1, 2]])


def _mask(s):
    return np.ascontiguousarray(MASK[s]).ctypes.data_as(IP)


def _c(a):
    return D(np.ascontiguousarray(a))


def coarse_partner(k, s, nb=None, oth=None):
    S, psi = np.empty(6), np.empty(19)
    L.acpo_strain_coarse_partner(ctypes.byref(k), DX, _mask(s), _c(NOW[s] if nb is None else nb),
                                 _c(FINE[s] if oth is None else oth), D(S), D(psi))
    return S, psi


def half_partner(k, s, pre=None):
    S, psi = np.empty(6), np.empty(19)
    L.acpo_strain_half_coarse_partner(ctypes.byref(k), DX, _mask(s), _c(PRE[s] if pre is None else pre),
                                      _c(NOW[s]), _c(FINE[s]), D(S), D(psi))
    return S, psi


def fine_partner(k, s, spacing=DX / 2.0):
    S, psi = np.empty(6), np.empty(19)
    L.acpo_strain_fine_partner(ctypes.byref(k), spacing, _c(FINE[s]), D(S), D(psi))
    return S, psi


def direct(k, s, level, spacing=None):
    S, psi = np.empty(6), np.empty(19)
    sp = (DX if level == 0 else DX / 2.0) if spacing is None else spacing
    L.acpo_strain_direct_coupling(ctypes.byref(k), level, sp, _mask(s), _c(PRE[s]), _c(NOW[s]), _c(FINE[s]),
                                  D(S), D(psi))
    return S, psi


def _ref(name, s):
    return K[name].reshape(NS, -1)[s]


def test_fixture_covers_both_kinds_of_neighbour():
    assert NS >= 32
    frac = MASK.mean()
    assert 0.3 < frac < 0.7  # grid-coupling / missing and ordinary neighbours both occur


@pytest.mark.parametrize("which", ["full", "half", "fine", "dc0", "dc1"])
def test_partner_functions_match_the_reference_bitexact(which):
    k = _consts()
    for s in range(NS):
        if which == "full":
            S, psi = coarse_partner(k, s)
        elif which == "half":
            S, psi = half_partner(k, s)
        elif which == "fine":
            S, psi = fine_partner(k, s)
        elif which == "dc0":
            S, psi = direct(k, s, 0)
        else:
            S, psi = direct(k, s, 1)
        assert np.array_equal(S, _s6_from_ref(_ref(f"{which}_s", s))), (which, s)
        assert np.array_equal(psi, _ref(f"{which}_psi", s)), (which, s)


def test_direct_coupling_on_level_0_sets_the_partner():
    """The level-0 direct coupling hands its strain and cubic terms to the fine
    partner (Node.cpp:2065-2096): the partner's values equal the node's."""
    for s in range(NS):
        assert np.array_equal(_ref("dc0p_s", s), _ref("dc0_s", s))
        assert np.array_equal(_ref("dc0p_psi", s), _ref("dc0_psi", s))


def test_planted_mistakes_are_caught():
    k = _consts()
    s = next(i for i in range(NS) if 0 < MASK[i].sum() < 6)
    ref_full = _s6_from_ref(_ref("full_s", s))
    # the other level-0 array instead of the current one
    assert not np.array_equal(coarse_partner(k, s, nb=OTHER[s])[0], ref_full)
    # the fine partner's other array
    assert not np.array_equal(coarse_partner(k, s, oth=FINE_OTHER[s])[0], ref_full)
    # no time average at t + 1/2: the current state in place of the pre-collision one
    assert not np.array_equal(half_partner(k, s, pre=NOW[s])[0], _s6_from_ref(_ref("half_s", s)))
    # a fine stencil on level 1 where the reference takes the coarse one
    assert not np.array_equal(direct(k, s, 1, spacing=DX / 4.0)[0], _s6_from_ref(_ref("dc1_s", s)))
    # the coarse spacing for the fine partners
    assert not np.array_equal(fine_partner(k, s, spacing=DX)[0], _s6_from_ref(_ref("fine_s", s)))
