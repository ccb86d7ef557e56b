"""Copy data files of the reference into tests/golden/ (unchanged):

* thermo/NCHO_photo_network.txt (the network of the reference's HD 189733b
  configuration) -> tests/golden/networks/;
* thermo/all_compose.txt (atoms and masses per species) -> tests/golden/;
* thermo/NASA9/<species>.txt for every species of that network, plus a few
  used by the Gibbs-energy pins -> tests/golden/NASA9/.

Run once in the build container:  python tests/golden/copy_reference_data.py
"""
import shutil
import sys
from pathlib import Path

SRC = Path("/root/reference/thermo")
DST = Path(__file__).resolve().parent
NETWORKS = ("NCHO_photo_network.txt",)
PIN_SPECIES = ("H", "H2", "O", "O2", "OH", "H2O", "CO", "CO2", "CH3", "CH4")


def _copy(src: Path, dst: Path):
    dst.parent.mkdir(parents=True, exist_ok=True)
    shutil.copyfile(src, dst)
    dst.chmod(0o644)


if __name__ == "__main__":
    sys.path.insert(0, str(DST.parent.parent / "jax-vulcan_amd"))
    from vulcan_amd.network import parse_network
    _copy(SRC / "all_compose.txt", DST / "all_compose.txt")
    species = set(PIN_SPECIES)
    for name in NETWORKS:
        _copy(SRC / name, DST / "networks" / name)
        species |= set(parse_network(SRC / name, extra_species=()).species)
    for sp in sorted(species):
        f = SRC / "NASA9" / f"{sp}.txt"
        if f.exists():
            _copy(f, DST / "NASA9" / f"{sp}.txt")
