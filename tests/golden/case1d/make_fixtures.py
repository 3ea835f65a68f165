"""Copies the data files of the reference's 1-D test cases that the tests
use into this folder (run in a checkout that has the reference at
/root/reference; the copies travel with the tests):

* ../input1d/<case>.txt: every 1-D case's input.txt (the parser fixtures);
* 1dB002HarmBar/: the case inputs (input.txt, bath.dat, gauges.dat,
  influx.dat) and the outputs the reference's authors ran it to
  (output/influx.dat, output/Hawassi_eta_gauge.txt, output/Constants_01.dat), the
  large ones gzipped;
* 1dFBr005Bor/: the case inputs (input.txt, gauges.dat, initial.dat),
  Output/Hawassi_eta_gauge.txt, Output/breaking.dat and Output/Constants_01.dat;
* 1dBBr002IrBar/: the case inputs, output/breaking.dat and Constants_01.dat;
* 1dB001IrSlope/: the case inputs, output/Output/influx.dat and Constants_01.dat;
* 1dBBr001HarmBar/: the case inputs (a 2000-node grid; no outputs ship).
"""
import glob
import gzip
import os
import shutil

REF = "/root/reference/1d/Testcases"
HERE = os.path.dirname(os.path.abspath(__file__))


def put(src, dst, gz=False):
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    if gz:
        with open(src, "rb") as f, gzip.GzipFile(dst + ".gz", "wb", mtime=0) as g:
            g.write(f.read())
    else:
        shutil.copyfile(src, dst)


def main():
    for f in sorted(glob.glob(f"{REF}/*/input.txt") + glob.glob(f"{REF}/*/Input/input.txt")):
        rel = os.path.relpath(f, REF).split(os.sep)[0]
        put(f, os.path.join(HERE, "..", "input1d", rel + ".txt"))
    c = os.path.join(HERE, "1dB002HarmBar")
    for name in ("input.txt", "bath.dat", "gauges.dat"):
        put(f"{REF}/1dB002HarmBar/{name}", os.path.join(c, name))
    put(f"{REF}/1dB002HarmBar/influx.dat", os.path.join(c, "influx.dat"), gz=True)
    put(f"{REF}/1dB002HarmBar/output/influx.dat", os.path.join(c, "expected", "influx.dat"), gz=True)
    put(f"{REF}/1dB002HarmBar/output/Hawassi_eta_gauge.txt", os.path.join(c, "expected", "Hawassi_eta_gauge.txt"),
        gz=True)
    put(f"{REF}/1dB002HarmBar/output/Constants_01.dat", os.path.join(c, "expected", "Constants_01.dat"), gz=True)
    c = os.path.join(HERE, "1dFBr005Bor")
    for name in ("input.txt", "gauges.dat"):
        put(f"{REF}/1dFBr005Bor/Input/{name}", os.path.join(c, name))
    put(f"{REF}/1dFBr005Bor/Input/initial.dat", os.path.join(c, "initial.dat"), gz=True)
    put(f"{REF}/1dFBr005Bor/Output/Hawassi_eta_gauge.txt", os.path.join(c, "expected", "Hawassi_eta_gauge.txt"),
        gz=True)
    put(f"{REF}/1dFBr005Bor/Output/breaking.dat", os.path.join(c, "expected", "breaking.dat"), gz=True)
    put(f"{REF}/1dFBr005Bor/Output/Constants_01.dat", os.path.join(c, "expected", "Constants_01.dat"), gz=True)
    c = os.path.join(HERE, "1dBBr002IrBar")
    for name in ("input.txt", "bath.dat", "gauges.dat"):
        put(f"{REF}/1dBBr002IrBar/{name}", os.path.join(c, name))
    put(f"{REF}/1dBBr002IrBar/influx.dat", os.path.join(c, "influx.dat"), gz=True)
    put(f"{REF}/1dBBr002IrBar/output/breaking.dat", os.path.join(c, "expected", "breaking.dat"), gz=True)
    put(f"{REF}/1dBBr002IrBar/output/Constants_01.dat", os.path.join(c, "expected", "Constants_01.dat"), gz=True)
    c = os.path.join(HERE, "1dB001IrSlope")
    for name in ("input.txt", "bath.dat", "gauges.dat"):
        put(f"{REF}/1dB001IrSlope/Input/{name}", os.path.join(c, name))
    put(f"{REF}/1dB001IrSlope/Input/influx.dat", os.path.join(c, "influx.dat"), gz=True)
    put(f"{REF}/1dB001IrSlope/output/Output/influx.dat", os.path.join(c, "expected", "influx.dat"), gz=True)
    put(f"{REF}/1dB001IrSlope/output/Output/Constants_01.dat", os.path.join(c, "expected", "Constants_01.dat"),
        gz=True)
    c = os.path.join(HERE, "1dBBr001HarmBar")
    for name in ("input.txt", "bath.dat", "gauges.dat"):
        put(f"{REF}/1dBBr001HarmBar/{name}", os.path.join(c, name))
    put(f"{REF}/1dBBr001HarmBar/influx.dat", os.path.join(c, "influx.dat"), gz=True)


if __name__ == "__main__":
    main()
