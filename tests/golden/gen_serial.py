"""Writes the serialisation fixtures: a ciphertext document and a key document
in the reference's layouts, built here from the reference's rules with
Python string formatting (no part of the product or the oracle is used):

* a polynomial is poly_export's text (aoa-ntt/src/arithmetic/poly_t.cpp:426-447):
  every residue of its base, limb after limb, in the coefficient domain, as a
  decimal number followed by a comma;
* a ciphertext is cipher_export's document (aoa-ntt/src/ckks/cipher_t.cpp:42-68):
  {"ct": {"0": c0, "1": c1}, "props": {"level", "scale", "slots"}} with the
  members in that order;
* the keys are CKKSContext::export_keys's document
  (aoa-ntt/src/ckks/ckkscontext.cpp:60-100): {"sk": {"s"}, "pk": {"a", "b"},
  "evk": {"a", "b"}}, each key a QB polynomial (Q limbs then special limbs,
  aoa-ntt/src/ckks/ckkskeys.cpp:12-15, :63, :83-84);
* the documents are written as RapidJSON's Writer writes them: no whitespace.

Parameters: N=2048, k=2 (q0 and one 55-bit prime), kl=1 (the next 55-bit
prime), prec 55, as the reference selects them (manager.cu:426-464); the
primes are read from aoa_amd/csrc/prime_table.inc.  Residues: one in 32
uniform in [0, q) (random.Random(2048)), q - 1 and 0 at fixed places, small
values elsewhere (so that the files stay small); every residue is below its
prime.

Run from the repository root: python tests/golden/gen_serial.py
"""
import gzip
import os
import random
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
N, K_Q, K_B, PREC = 2048, 2, 1, 55
LEVEL, SCALE, SLOTS = 1, 1 << 55, 1024
CT_FILE = os.path.join(HERE, "serial_n2048_ct.json.gz")
KEYS_FILE = os.path.join(HERE, "serial_n2048_keys.json.gz")


def primes():
    """q0 (the first 63-bit prime), then the first K_Q - 1 + K_B primes of the
    55-bit bucket (prec 55 continues that bucket for the special primes)."""
    text = open(os.path.join(ROOT, "aoa_amd", "csrc", "prime_table.inc")).read()

    def bucket(bits):
        m = re.search(r"AOA_BUCKET_%d\[\d+\] = \{(.*?)\};" % bits, text, re.S)
        return [int(p) for p, _ in re.findall(r"\{(\d+)ULL, (\d+)\}", m.group(1))]
    b55 = bucket(55)
    return [bucket(63)[0]] + b55[:K_Q - 1 + K_B]


def poly(rng, ps):
    """Residues of one polynomial over the primes ps, limb-major."""
    out = []
    for r, q in enumerate(ps):
        for n in range(N):
            if n % 32 == 0:
                v = rng.randrange(q)
            elif n == 1:
                v = q - 1
            elif n == 2:
                v = 0
            else:
                v = (n * 31 + r * 7) % 1000
            out.append(v)
    return out


def poly_text(res):
    return "".join("%d," % v for v in res)


def documents():
    """(ciphertext document, key document, their residues)."""
    ps = primes()
    rng = random.Random(2048)
    q = ps[:K_Q]
    c0, c1 = poly(rng, q), poly(rng, q)
    ct = ('{"ct":{"0":"%s","1":"%s"},"props":{"level":%d,"scale":%d,"slots":%d}}'
          % (poly_text(c0), poly_text(c1), LEVEL, SCALE, SLOTS))
    keys = [poly(rng, ps) for _ in range(5)]  # sk.s, pk.a, pk.b, evk.a, evk.b
    kd = ('{"sk":{"s":"%s"},"pk":{"a":"%s","b":"%s"},"evk":{"a":"%s","b":"%s"}}'
          % tuple(poly_text(k) for k in keys))
    return ct, kd, {"ct": [c0, c1], "keys": keys, "primes": ps}


def main():
    ct, kd, _ = documents()
    for path, doc in ((CT_FILE, ct), (KEYS_FILE, kd)):
        with gzip.GzipFile(path, "wb", mtime=0) as f:
            f.write(doc.encode())
        print("wrote", path, len(doc), "bytes of text")


if __name__ == "__main__":
    main()
