"""Copies the reference's data fixtures into tests/golden/ (run where a checkout of the reference
exists; the tests and the GPU runs read only the copies)."""
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")

if __name__ == "__main__":
    shutil.copyfile(REF / "public" / "drape.bin", HERE / "drape_default.bin")
    print("copied", HERE / "drape_default.bin")
