"""Copy the NASA-9 fits of ten species from the reference's thermo/NASA9
data folder into tests/golden/NASA9/ (data files, unchanged).

Run once in the build container:  python tests/golden/copy_nasa9.py
"""
import shutil
from pathlib import Path

SRC = Path("/root/reference/thermo/NASA9")
DST = Path(__file__).resolve().parent / "NASA9"
SPECIES = ("H", "H2", "O", "O2", "OH", "H2O", "CO", "CO2", "CH3", "CH4")

if __name__ == "__main__":
    DST.mkdir(exist_ok=True)
    for sp in SPECIES:
        shutil.copyfile(SRC / f"{sp}.txt", DST / f"{sp}.txt")
        (DST / f"{sp}.txt").chmod(0o644)
