"""Write the tuning table the library loads (tuning_gfx950.txt) from the
"autotune" field of bench.py's JSON lines (and of earlier tables).

    python omega_h_v1_amd/make_tuning_table.py out.txt [old.txt] run1.json ...

Each line of the table is "tuner:2^class=variant", the online tuner's own
report format.  Where runs disagree on a size class, the run named first
wins; classes a run did not finish timing ("?") are left out, so the library
times them online.
"""
import json
import sys


def main(out, runs):
    table = {}
    for path in runs:
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith("{"):
                    entries = json.loads(line).get("autotune", [])
                elif ":2^" in line and not line.startswith("#"):
                    entries = [line]
                else:
                    continue
                for entry in entries:
                    key, _, v = entry.rpartition("=")
                    if v.isdigit() and key not in table:
                        table[key] = v
    with open(out, "w") as f:
        f.write("# tuner:2^size_class=variant, measured on MI355X (gfx950) by bench.py\n")
        for key in sorted(table):
            f.write("%s=%s\n" % (key, table[key]))
    print("%d entries -> %s" % (len(table), out))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2:])
