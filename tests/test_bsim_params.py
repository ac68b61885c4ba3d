"""build_bsim_params and parse_number (crates/sim-devices/src/bsim/mod.rs:
83-420): the reference's own tests (mod.rs:509-575) and the aliases, suffix
rules and number grammar the restatement follows."""
import math

import numpy as np
import pytest

from rustspice_amd import netlists as NL


def test_build_params_empty():  # mod.rs:513-519
    rec, parsed = NL.build_bsim_params({}, 49, False)
    assert parsed == {"level": 49, "pmos": False}
    assert rec == NL.bsim_model(False)
    assert rec["pmos"] == 0 and rec["vth0"] == 0.7


def test_build_params_with_values():  # mod.rs:521-532
    rec, _ = NL.build_bsim_params({"vth0": "0.5", "u0": "400", "tox": "2n"}, 49, False)
    assert abs(rec["vth0"] - 0.5) < 0.001
    assert abs(rec["u0"] - 400.0) < 0.1
    assert abs(rec["tox"] - 2e-9) < 1e-12


def test_parse_number_suffixes():  # mod.rs:534-541
    assert abs(NL.parse_number("1.5") - 1.5) < 1e-10
    assert abs(NL.parse_number("1n") - 1e-9) < 1e-15
    assert abs(NL.parse_number("1u") - 1e-6) < 1e-12
    assert abs(NL.parse_number("10k") - 1e4) < 1e-6
    assert abs(NL.parse_number("2.5meg") - 2.5e6) < 1.0


def test_build_params_bsim4():  # mod.rs:559-575
    rec, parsed = NL.build_bsim_params({"alpha0": "1e-6", "beta0": "30", "ku0": "0.1", "kvth0": "0.01",
                                        "jtss": "1e-10"}, 54, False)
    assert parsed["level"] == 54
    assert abs(rec["alpha0"] - 1e-6) < 1e-12
    assert abs(rec["beta0"] - 30.0) < 0.1
    assert abs(rec["ku0"] - 0.1) < 0.001
    assert abs(rec["kvth0"] - 0.01) < 0.001
    assert abs(rec["jtss"] - 1e-10) < 1e-15


@pytest.mark.parametrize("text,want", [
    ("1meg", 1e6), ("1MEG", 1e6), ("2mil", 2 * 25.4e-6), ("3f", 3 * 1e-15), ("3p", 3 * 1e-12),
    ("-2.5n", -2.5 * 1e-9), ("7m", 7 * 1e-3), ("2g", 2 * 1e9), ("3t", 3 * 1e12), ("1e3k", 1e3 * 1e3),
    ("5.", 5.0), (".5u", 0.5 * 1e-6), ("+3", 3.0), (" 4k ", 4e3), ("1e-3m", 1e-3 * 1e-3), ("12", 12.0),
])
def test_suffixes_multiply_the_number(text, want):
    """Each suffix is one multiplication of the parsed number (so the result
    is the product rounded once, as in the reference)."""
    assert NL.parse_number(text) == want


@pytest.mark.parametrize("text", ["", "m", "abc", "1 k", "1_0", "0x10", "1e", "e5", "1.2.3", "1ms", "--1"])
def test_what_does_not_parse(text):
    """Rust's f64 grammar: no inner whitespace, underscores or hex (Python's
    float() would take several of these)."""
    assert NL.parse_number(text) is None


def test_special_values_parse_as_rust_does():
    assert NL.parse_number("inf") == math.inf
    assert NL.parse_number("-infinity") == -math.inf
    assert math.isnan(NL.parse_number("nan"))
    # a one-letter suffix whose number part fails: the whole string is tried
    assert NL.parse_number("5e") is None


def test_aliases_in_order():
    rec, _ = NL.build_bsim_params({"vto": "0.45"}, 49, False)
    assert rec["vth0"] == 0.45
    rec, _ = NL.build_bsim_params({"vth": "0.3", "vto": "0.45"}, 49, False)
    assert rec["vth0"] == 0.45  # vto is tried before vth
    rec, _ = NL.build_bsim_params({"vth0": "0.2", "vto": "0.45"}, 49, False)
    assert rec["vth0"] == 0.2
    rec, _ = NL.build_bsim_params({"vth0": "bad", "vto": "0.45"}, 49, False)
    assert rec["vth0"] == 0.45  # an unparsable value falls through to the next alias
    rec, _ = NL.build_bsim_params({"uo": "350"}, 49, True)
    assert rec["u0"] == 350.0 and rec["pmos"] == 1 and rec["vth0"] == -0.7


def test_keys_are_looked_up_lower_case_only():
    """The reference lower-cases the key it looks for, not the map's keys."""
    rec, parsed = NL.build_bsim_params({"VTH0": "0.1", "U0": "100"}, 49, False)
    assert rec == NL.bsim_model(False) and "vth0" not in parsed


def test_tnom_and_pemod_conversions():
    rec, parsed = NL.build_bsim_params({"tnom": "27", "pemod": "2.7"}, 54, False)
    assert rec["tnom"] == 27.0 + 273.15
    assert parsed["pemod"] == 2
    _, parsed = NL.build_bsim_params({"pemod": "-1"}, 54, False)
    assert parsed["pemod"] == 0
    _, parsed = NL.build_bsim_params({"pemod": "1e12"}, 54, False)
    assert parsed["pemod"] == 4294967295


def test_every_parsed_field_reaches_the_record_or_is_kept():
    """Fields the device evaluation reads land in the rs_bsim_model record;
    the others (capacitances, BSIM4 extras) stay in `parsed`."""
    params = {k: "1.25" for field, keys in NL.BSIM_PARAM_KEYS for k in keys[:1]}
    rec, parsed = NL.build_bsim_params(params, 54, False)
    for field, _ in NL.BSIM_PARAM_KEYS:
        want = 1.25 + 273.15 if field == "tnom" else (1 if field == "pemod" else 1.25)
        assert parsed[field] == want
        if field in NL.BSIM_FIELDS:
            assert rec[field] == want
    untouched = [f for f in NL.BSIM_FIELDS if f not in parsed]
    base = NL.bsim_model(False)
    for f in untouched:
        assert rec[f] == base[f] or (np.isnan(rec[f]) and np.isnan(base[f]))
