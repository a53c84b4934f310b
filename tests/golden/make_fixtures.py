"""Convert the reference's golden final-pressure fields to raw fixtures.

Reads the three BSON.jl documents under the reference checkout's
``test/references/`` (used by test/test2D_xPU.jl:130,155,194) as DATA: a
small BSON walker extracts the ``pend`` array (Julia Float64, column-major)
and writes ``<name>.f64`` (little-endian float64, column-major, x fastest)
plus ``fixtures.json`` with shapes, the setup each field was produced with
(restated from test/test2D_xPU.jl:108-195) and checksums.

Nothing from the reference is executed.  Run once in a container that has the
reference checkout; the outputs are committed next to this script.

    python tests/golden/make_fixtures.py /root/reference
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _cstring(buf, pos):
    end = buf.index(b"\x00", pos)
    return buf[pos:end].decode("utf-8"), end + 1


def parse_bson(buf, pos=0):
    """Minimal BSON document walker (types used by BSON.jl arrays)."""
    (length,) = struct.unpack_from("<i", buf, pos)
    end = pos + length - 1  # trailing NUL
    pos += 4
    out = {}
    while pos < end:
        etype = buf[pos]
        pos += 1
        key, pos = _cstring(buf, pos)
        if etype == 0x01:  # double
            (out[key],) = struct.unpack_from("<d", buf, pos)
            pos += 8
        elif etype == 0x02:  # string
            (slen,) = struct.unpack_from("<i", buf, pos)
            out[key] = buf[pos + 4 : pos + 4 + slen - 1].decode("utf-8")
            pos += 4 + slen
        elif etype in (0x03, 0x04):  # document / array
            (dlen,) = struct.unpack_from("<i", buf, pos)
            sub = parse_bson(buf, pos)
            if etype == 0x04:
                sub = [sub[k] for k in sorted(sub, key=int)]
            out[key] = sub
            pos += dlen
        elif etype == 0x05:  # binary
            (blen,) = struct.unpack_from("<i", buf, pos)
            out[key] = bytes(buf[pos + 5 : pos + 5 + blen])
            pos += 5 + blen
        elif etype == 0x08:  # bool
            out[key] = bool(buf[pos])
            pos += 1
        elif etype == 0x10:  # int32
            (out[key],) = struct.unpack_from("<i", buf, pos)
            pos += 4
        elif etype == 0x12:  # int64
            (out[key],) = struct.unpack_from("<q", buf, pos)
            pos += 8
        elif etype == 0x0A:  # null
            out[key] = None
        else:
            raise ValueError(f"unsupported BSON element type 0x{etype:02x}")
    return out


def julia_array(doc):
    assert doc["tag"] == "array", doc.get("tag")
    assert doc["type"]["name"] == ["Core", "Float64"], doc["type"]
    size = [int(s) for s in doc["size"]]
    data = np.frombuffer(doc["data"], dtype="<f8")
    assert data.size == int(np.prod(size)), (data.size, size)
    # column-major: reshape with Fortran order -> numpy array indexed [i, j]
    return data.reshape(size, order="F"), size


# Setup of each fixture, restated from test/test2D_xPU.jl (all positions in
# meters, rows = items).  "vel" is a recipe, built by the tests.
FIXTURES = {
    "acoustic2D_center_halo20": {
        "cite": "test/test2D_xPU.jl:108-131",
        "n": [201, 201], "l": [2000.0, 2000.0], "lt": 2.0,
        "vel": {"kind": "constant", "c": 2000.0},
        "f0": 10.0, "t0": 0.4,
        "srcpos": [[1000.0, 1000.0]],
        "recpos": [[1000.0, 2 * 2000.0 / 3], [2 * 2000.0 / 3, 2 * 2000.0 / 3]],
        "halo": 20, "rcoef": 0.0001, "freetop": False,
    },
    "acoustic2D_center_freetop_halo20": {
        "cite": "test/test2D_xPU.jl:133-156",
        "n": [201, 201], "l": [2000.0, 2000.0], "lt": 2.0,
        "vel": {"kind": "constant", "c": 2000.0},
        "f0": 10.0, "t0": 0.4,
        "srcpos": [[1000.0, 1000.0]],
        "recpos": [[1000.0, 2 * 2000.0 / 3], [2 * 2000.0 / 3, 2 * 2000.0 / 3]],
        "halo": 20, "rcoef": 0.0001, "freetop": True,
    },
    "acoustic2D_gradient_freetop_halo20": {
        "cite": "test/test2D_xPU.jl:158-195",
        "n": [422, 242], "l": [2100.0 * 5, 1200.0 * 5], "lt": 5.0,
        "vel": {"kind": "gradient", "v0": 2000.0, "dv": 12.0, "const_from_end": 40},
        "f0": 10.0, "t0": 0.4,
        "srcpos": [[k * 2100.0 * 5 / 11, 5.0] for k in (3, 4, 5, 6, 7, 8)],
        "recpos": [[2100.0 * 5 / 2, 2 * 1200.0 * 5 / 3], [2 * 2100.0 * 5 / 3, 2 * 1200.0 * 5 / 3]],
        "halo": 20, "rcoef": 0.0001, "freetop": True,
    },
}


def main(ref_root):
    meta = {}
    for name, setup in FIXTURES.items():
        path = os.path.join(ref_root, "test", "references", name + ".bson")
        with open(path, "rb") as fh:
            doc = parse_bson(fh.read())
        pend, size = julia_array(doc["pend"])
        assert size == setup["n"], (name, size, setup["n"])
        raw = np.asfortranarray(pend).ravel(order="F").astype("<f8").tobytes()
        with open(os.path.join(HERE, name + ".f64"), "wb") as fh:
            fh.write(raw)
        meta[name] = dict(setup)
        meta[name].update({
            "file": name + ".f64",
            "dtype": "float64-le",
            "order": "column-major (x fastest)",
            "shape": size,
            "sha256": hashlib.sha256(raw).hexdigest(),
            "max_abs": float(np.abs(pend).max()),
            "l2": float(np.sqrt((pend.astype(np.float64) ** 2).sum())),
            "source": "test/references/" + name + ".bson (pend)",
        })
        print(name, size, meta[name]["max_abs"], meta[name]["l2"])
    with open(os.path.join(HERE, "fixtures.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
