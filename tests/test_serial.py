"""The serialisation fixtures (tests/golden/serial_n2048_*.json.gz) on the CPU:
they are what gen_serial.py writes from the reference's format rules, they
parse as JSON with the reference's members in the reference's order, and
every residue list has the length of its base and values below its primes."""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import gen_serial  # noqa: E402


def _text(path):
    return gzip.open(path).read().decode()


def _residues(s):
    assert s.endswith(","), "poly_export ends every residue with a comma"
    return [int(x) for x in s[:-1].split(",")]


def test_fixtures_are_the_generators_output():
    ct, kd, _ = gen_serial.documents()
    assert _text(gen_serial.CT_FILE) == ct
    assert _text(gen_serial.KEYS_FILE) == kd


def test_ciphertext_document_layout():
    doc = _text(gen_serial.CT_FILE)
    assert " " not in doc and "\n" not in doc  # the Writer's compact text
    j = json.loads(doc)
    assert list(j) == ["ct", "props"] and list(j["ct"]) == ["0", "1"]
    assert list(j["props"]) == ["level", "scale", "slots"]
    assert j["props"] == {"level": gen_serial.LEVEL, "scale": gen_serial.SCALE, "slots": gen_serial.SLOTS}
    ps = gen_serial.primes()
    for y in "01":
        r = _residues(j["ct"][y])
        assert len(r) == gen_serial.K_Q * gen_serial.N
        for limb in range(gen_serial.K_Q):
            seg = r[limb * gen_serial.N:(limb + 1) * gen_serial.N]
            assert max(seg) == ps[limb] - 1 and min(seg) == 0


def test_key_document_layout():
    j = json.loads(_text(gen_serial.KEYS_FILE))
    assert list(j) == ["sk", "pk", "evk"]
    assert list(j["sk"]) == ["s"] and list(j["pk"]) == ["a", "b"] and list(j["evk"]) == ["a", "b"]
    ps = gen_serial.primes()
    R = gen_serial.K_Q + gen_serial.K_B
    for grp, mem in (("sk", "s"), ("pk", "a"), ("pk", "b"), ("evk", "a"), ("evk", "b")):
        r = _residues(j[grp][mem])
        assert len(r) == R * gen_serial.N, (grp, mem)
        for limb in range(R):
            assert max(r[limb * gen_serial.N:(limb + 1) * gen_serial.N]) < ps[limb]


def test_prime_layout_matches_the_oracle():
    from orc import Oracle
    o = Oracle(gen_serial.N, gen_serial.K_Q, gen_serial.K_B, gen_serial.PREC)
    assert gen_serial.primes() == list(o.primes)
