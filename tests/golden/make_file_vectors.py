"""Writes tests/golden/file_vectors.json from the reference's own file readers
and writers (oracle/_ref/libxmcml_ref_io.so, built by `make -C oracle ref`
from the reference sources): the parsed values of every input fixture under
tests/golden/files/, and SHA-256 digests of the .mcml.out and .mcml.bk files
the reference writes for a seeded output of the quirks input.

    python tests/golden/make_file_vectors.py
"""
import json
import os
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import files_lib as fl  # noqa: E402
from xmcml_amd.files import input_to_spec  # noqa: E402

INPUTS = {  # name: (xml, surface or None)
    "quirks": ("quirks.xml", None),
    "diffuse_830_ring": ("diffuse_830_ring.xml", None),
    "cosine_exp": ("cosine_exp.xml", "cosine_exp.surface"),
    "stripes_exp": ("stripes_exp.xml", "stripes_exp.surface"),
}
SEED = 20240607


def main():
    out = {"parsed": {}, "writers": {}}
    for name, (xml, surf) in INPUTS.items():
        r = fl.ref_parse(fl.FIXTURES / xml, None if surf is None else fl.FIXTURES / surf)
        assert r is not None, name
        out["parsed"][name] = fl.canonical(*r)
    x, b = fl.ref_parse(fl.FIXTURES / "quirks.xml")
    spec = input_to_spec(x)
    o = fl.filled_output(spec, SEED)
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            assert fl.ref_io().refio_write_output(fl.C.byref(x), o.ptr, b"result.mcml.out") == 1
            assert fl.ref_io().refio_write_backup(fl.C.byref(x), o.ptr, 777, 123456789) == 1
            bk = [f for f in os.listdir(".") if f.endswith(".mcml.bk")]
            assert len(bk) == 1
            out["writers"] = {"seed": SEED, "output_sha256": fl.sha256("result.mcml.out"),
                              "backup_name": bk[0], "backup_sha256": fl.sha256(bk[0]),
                              "rng": [777, 123456789]}
        finally:
            os.chdir(cwd)
    (HERE / "file_vectors.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
