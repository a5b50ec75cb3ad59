"""Copies the reference's own test inputs into tests/golden/ (run in a container that holds
the reference at /root/reference).  They are data files its harnesses read:

  c++/datasets/cluster.dat.all  800 x 300 CSV, the input of c++/src/test/gpu_unit_test.cu:115-133
                                -> cluster.dat.all.gz (gzip, the same bytes)
  c++/datasets/small.csv        2 x 3 CSV, the input of the test_multicore target
                                (c++/multicore.mk:18-21) -> small.csv

SHA-256 of the uncompressed files is recorded in reference_inputs.json, which the tests check.
"""
import gzip
import hashlib
import json
import os
import shutil

SRC = "/root/reference/c++/datasets"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    meta = {}
    for name, out, packed in (("cluster.dat.all", "cluster.dat.all.gz", True), ("small.csv", "small.csv", False)):
        src = os.path.join(SRC, name)
        data = open(src, "rb").read()
        dst = os.path.join(HERE, out)
        if packed:
            with gzip.GzipFile(dst, "wb", compresslevel=9, mtime=0) as f:
                f.write(data)
        else:
            shutil.copyfile(src, dst)
        meta[out] = {"source": "c++/datasets/" + name, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    with open(os.path.join(HERE, "reference_inputs.json"), "w") as f:
        json.dump(meta, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
