"""Writes tests/golden/edga_result_head.dat: the header and the first 300
determinant lines of the reference's EDGA example trial
(examples/dmrg_afqmc/edga_result.dat, 11857 determinants over 10 active
orbitals, 5 alpha + 5 beta electrons).  Data only; run from the repository
root where /root/reference is present."""
import pathlib

SRC = pathlib.Path('/root/reference/examples/dmrg_afqmc/edga_result.dat')
DST = pathlib.Path(__file__).with_name('edga_result_head.dat')
NDET = 300

out, n = [], 0
for line in SRC.read_text().splitlines():
    if line.lstrip().startswith('#'):
        out.append(line)
        continue
    if n < NDET and line.strip():
        out.append(line)
        n += 1
DST.write_text('\n'.join(out) + '\n')
print(f"wrote {DST} ({n} determinants)")
