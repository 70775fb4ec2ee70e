"""Regenerate the golden fixtures of the parity scenarios from the reference itself.

Runs only where the reference sources are present: oracle/Makefile compiles them (with the
reference's own flags) into oracle/_ref/libref.a and links tests/golden/drive.c against it
(oracle/_ref/drive_ref).  Each scenario's output lines are written to
tests/golden/<scenario>.txt, the text files of the scenarios in TEXT (alis_drive_text: rows,
Poynting and far-field spectra, far-field cuts) to tests/golden/text/<scenario>/, and
manifest.json records the step counts and a SHA-256 of the reference sources the fixtures
came from.

    python tests/golden/make_golden.py [SCENARIO ...]     (default: all)
"""
import hashlib
import json
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path("/root/reference/src")

# scenario -> steps (each finishes in well under a second on the reference)
SCENARIOS = {
    "dipole": 200,
    "symband": 200,
    "nanoblock": 200,
    "oblique": 200,
    "periodic": 200,
    "hexagonal": 200,
    "multipole": 200,
    "pillars": 200,
    # 18 dispersive spheres (gold and silver) lit by a TF/SF plane wave
    "particles": 200,
    # a pump and a signal world coupled every step by nonLinearSHG (second-order d tensor)
    "shg": 200,
    # 40 Bloch k-point worlds of a photonic-crystal slab (kp0..kp39), stepped together on the
    # GPU (80 batch entries): probes and writeSpectrum files
    "kpoints": 200,
}

# scenario -> steps for the text-output fixtures (tests/test_text.py)
TEXT = {"symband": 200, "nanoblock": 200}


def main():
    if not (REF / "update.c").exists():
        sys.exit("reference sources not found; fixtures can only be regenerated where they are")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref", "ref-programs"], check=True)
    drive = ROOT / "oracle" / "_ref" / "drive_ref"
    only = sys.argv[1:]
    for name, steps in SCENARIOS.items():
        if only and name not in only:
            continue
        out = HERE / f"{name}.txt"
        with tempfile.TemporaryDirectory() as d:
            # scenarios write their spectrum files next to the output file
            tmp = Path(d) / "out.txt"
            subprocess.run([str(drive), name, str(steps), str(tmp)], check=True, stdout=subprocess.DEVNULL,
                           env={"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"}, cwd=d)
            shutil.copyfile(tmp, out)
        print(f"{name}: {sum(1 for _ in out.open())} lines")
    for name, steps in TEXT.items():
        if only and name not in only:
            continue
        out = HERE / "text" / name
        shutil.rmtree(out, ignore_errors=True)
        out.mkdir(parents=True)
        subprocess.run([str(drive), "--text", name, str(steps), str(out)], check=True, stdout=subprocess.DEVNULL,
                       env={"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"}, cwd=out)
        print(f"{name} text: {sorted(p.name for p in out.iterdir())}")
    h = hashlib.sha256()
    for f in sorted(REF.glob("*.[ch]")):
        h.update(f.name.encode())
        h.update(f.read_bytes())
    manifest = {"scenarios": SCENARIOS, "text": TEXT, "reference_sources_sha256": h.hexdigest(),
                "generator": "oracle/_ref/drive_ref (tests/golden/drive.c over the reference library)"}
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")


if __name__ == "__main__":
    main()
