"""Network constants on the host side (core/Constant.py:16-120): which model
names become constant slots, and the module / network copies of a constant."""

import pytest

import annarchy_amd as ah
from annarchy_amd import codegen
from annarchy_amd import network as ahnet


def _slots(equations, parameters="", spike=None, constants=("k_c",)):
    src = codegen.generate_from_text("n", parameters, equations, spike, None, None, constants=list(constants))
    return {s[0]: s for s in src.slots}


def test_constant_becomes_population_wide_slot():
    s = _slots("r = r + k_c")
    assert s["k_c"][1:] == (False, 0.0, 4)
    assert s["r"][3] == 1


def test_attribute_hides_constant():
    s = _slots("r = r + k_c", parameters="k_c = 0.5")
    assert s["k_c"][3] == 0 and s["k_c"][2] == 0.5


def test_only_whole_names_match():
    s = _slots("r = r + k_c2 + 1.0", parameters="k_c2 = 1.0")
    assert "k_c" not in s
    s = _slots("v = v + 1.0", spike="v > k_c")
    assert s["k_c"][3] == 4   # read in the spike condition only


def test_unused_constant_adds_nothing():
    a = codegen.generate_from_text("n", "", "r = r + 1.0", None, None, None, constants=["k_c"])
    b = codegen.generate_from_text("n", "", "r = r + 1.0", None, None, None)
    assert a.source == b.source


def test_synapse_constant_locality():
    src = codegen.generate_synapse_from_text("s", "", "w = w + eta_c * pre.k_c", constants=["eta_c", "k_c"])
    assert src.desc.constant_names() == ["eta_c"]   # pre.k_c is a neuron attribute
    assert "const int32_t kAttrLocality[] = {3};" in src.source


def test_constant_copies():
    c = ah.Constant("tmp_c", 2.0)
    try:
        assert float(c) == 2.0 and c.value == 2.0
        net = ah.Network()
        local = net.get_constant("tmp_c")
        assert local is not c and local.value == 2.0
        c.set(5.0)
        assert local.value == 5.0
        local.set(7.0)
        assert c.value == 5.0
        own = net.constant("own_c", 1.0)
        assert net.get_constant("own_c") is own
        with pytest.raises(KeyError):
            ah.Network().get_constant("own_c")
    finally:
        ahnet._GLOBAL_CONSTANTS.pop("tmp_c", None)


def test_slot_list_follows_projections_added_after_an_attribute_write():
    # setting an attribute before connect() must not freeze the slot list:
    # the projection adds the population's _sum_exc (core/Population.py:617-635)
    import warnings

    import annarchy_amd as ah

    net = ah.Network(dt=1.0)
    pre = net.create(4, ah.Neuron(equations="r = 1", name="SlotPre"))
    post = net.create(3, ah.Neuron(parameters="b = 0.0", equations="r = b", name="SlotPost"))
    post.b = 1.0
    assert "_sum_exc" not in post._slot_info()
    net.connect(pre, post, "exc")
    assert "_sum_exc" in post._slot_info()
    # a target no projection uses: zeros of the population's shape, with a warning
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        z = post.sum("inh")
    assert z.shape == (3,) and not z.any() and len(w) == 1
