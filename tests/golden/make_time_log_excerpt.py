"""Writes tests/golden/time_log_excerpt.txt: the header and the first 40
`time:` lines of the reference's energy_check/input/time.log (a run log it
ships as data), used to pin the time.log reader and the energy report."""
from pathlib import Path

SRC = Path("/root/reference/energy_check/input/time.log")
OUT = Path(__file__).with_name("time_log_excerpt.txt")

lines, n_time = [], 0
for line in SRC.open():
    if line.startswith("time:"):
        n_time += 1
        if n_time > 40:
            break
    lines.append(line)
OUT.write_text("".join(lines))
print(f"wrote {OUT} ({len(lines)} lines)")
