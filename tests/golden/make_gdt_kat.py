"""Write tests/golden/gdt_kat.json from the reference's own gdt headers.

oracle/_ref/ref_gdt is built by `make -C oracle ref` from
common/gdt/gdt/random/random.h and common/gdt/gdt/math/vec.h where they lie
in the reference checkout; this script runs it and stores its output (known
answers for LCG<16> seeding/draws and vec3f cross/dot/normalize/length).
"""
import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = HERE.parent.parent / "oracle" / "_ref" / "ref_gdt"

if __name__ == "__main__":
    out = subprocess.run([str(REF)], check=True, stdout=subprocess.PIPE, text=True).stdout
    data = json.loads(out)
    (HERE / "gdt_kat.json").write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {len(data['lcg'])} LCG seeds, {len(data['vec'])} vector cases")
