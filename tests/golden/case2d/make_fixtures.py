"""Copies the data files of the reference's 2-D test cases whose grids are not
powers of two into this folder (run in a checkout that has the reference at
/root/reference; the copies travel with the tests): each case's input.txt,
and bath.dat / influx.dat gzipped.

* 2dBBr001_HarmBr_BB93: 2000 x 2 (quasi-1-D flume), HS3, three reference depths
* 2dBBr002_Irreg:       300 x 250, HS3, two reference depths, breaking
* 2dFBr001_Irreg:       600 x 300, HS2, flat bottom, breaking
* 2dFBr001_TUDFoc7:     4096 x 2 (quasi-1-D flume), HS3, flat bottom, breaking
"""
import gzip
import os
import shutil

REF = "/root/reference/2d/Testcases"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ("2dBBr001_HarmBr_BB93", "2dBBr002_Irreg", "2dFBr001_Irreg", "2dFBr001_TUDFoc7")


def main():
    for case in CASES:
        d = os.path.join(HERE, case)
        os.makedirs(d, exist_ok=True)
        shutil.copyfile(os.path.join(REF, case, "input.txt"), os.path.join(d, "input.txt"))
        for name in ("bath.dat", "influx.dat"):
            src = os.path.join(REF, case, name)
            if os.path.exists(src):
                with open(src, "rb") as f, gzip.GzipFile(os.path.join(d, name + ".gz"), "wb", mtime=0) as g:
                    g.write(f.read())


def materialize(case: str, dst: str) -> str:
    """The case directory as the reference reads it (input.txt, bath.dat,
    influx.dat), decompressed into dst; returns dst."""
    src = os.path.join(HERE, case)
    os.makedirs(dst, exist_ok=True)
    for name in os.listdir(src):
        p = os.path.join(src, name)
        if name.endswith(".gz"):
            with gzip.open(p, "rb") as f, open(os.path.join(dst, name[:-3]), "wb") as g:
                g.write(f.read())
        else:
            shutil.copyfile(p, os.path.join(dst, name))
    return dst


if __name__ == "__main__":
    main()
