"""Copy the reference's 2-D spiral dataset (data/spiral.mat: k = kx + i ky in cycles
per sample, |k| <= 0.5, 1182 x 60; w = density weights; data = 1182 x 60 x 8 coil
samples; from M. Lustig's SparseMRI/ESPIRiT demos, see data/readme.txt) into
tests/golden/spiral.mat.  A data fixture: the file is copied unchanged.
Run from the repository root with the reference checked out at /root/reference:
    python tests/golden/make_spiral.py
"""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/data/spiral.mat"

if __name__ == "__main__":
    shutil.copyfile(SRC, os.path.join(HERE, "spiral.mat"))
    print("copied", SRC)
