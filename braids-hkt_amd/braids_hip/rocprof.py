"""Summaries of rocprofv3 CSV output for the hydraulic-step kernels.

Two inputs are read:
  * ``--kernel-trace --stats`` -> ``<prefix>_kernel_stats.csv`` (per-kernel
    call count and average duration);
  * two ``--pmc`` runs, one with ``FETCH_SIZE`` and one with ``WRITE_SIZE``
    -> ``<prefix>_counter_collection.csv`` (KiB per dispatch).

HBM bytes per dispatch are the counter readings scaled by what the same
counter reads for ``k_stream_copy`` (a float4 copy of a known byte count, run
in the same process): ``factor = known_bytes / counter_bytes``.  On gfx950
the copy of 1 GiB reads FETCH_SIZE = 512 MiB and WRITE_SIZE = 1 GiB, so the
fetch factor is 2 and the write factor 1.

    python -m braids_hip.rocprof --stats S.csv --fetch F.csv --write W.csv \
        --copy-bytes 1073741824 --title ... --command ... > summary.md
"""
import argparse
import collections
import csv
import re

# kernel-name prefix -> kind used by the solver's event profiler
KINDS = {'k_prep': 'prep', 'k_prep_ls64': 'prep', 'k_traces': 'traces', 'k_face': 'face', 'k_face_traces': 'face',
         'k_close': 'close', 'k_finalize_mirror': 'finalize',
         'k_face_f32': 'face', 'k_drain': 'drain', 'k_gather': 'gather', 'k_ledger': 'ledger',
         'k_river_fused': 'rfused', 'k_river_update': 'rupdate', 'k_river_drag': 'rdrag',
         'k_river_ledger': 'ledger', 'kh_reconstruct': 'hrecon', 'kh_faces': 'hface', 'kh_budget': 'hbudget',
         'kh_update': 'hupdate', 'kh_predict': 'hpredict', 'k_stream_copy': 'copy'}


def short_name(name):
    n = name.replace('(anonymous namespace)::', '').replace('void ', '')
    n = re.sub(r'\(.*$', '', n)
    return n.replace('braids::', '')


def kind_of(name):
    return KINDS.get(short_name(name).split('<')[0])


def read_stats(path):
    """{short kernel name: (calls, average ns)} from a kernel_stats.csv."""
    out = {}
    for r in csv.DictReader(open(path)):
        out[short_name(r['Name'])] = (int(r['Calls']), float(r['AverageNs']))
    return out


def read_counter(path, counter):
    """{short kernel name: [KiB per dispatch, ...]} from a counter_collection.csv."""
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r['Counter_Name'] == counter:
            d[short_name(r['Kernel_Name'])].append(float(r['Counter_Value']))
    return d


def traffic_table(fetch_csv, write_csv, copy_bytes):
    """Per-kernel mean HBM bytes per dispatch, calibrated on k_stream_copy.

    Returns (rows, fetch_factor, write_factor, copy_readings)."""
    f = read_counter(fetch_csv, 'FETCH_SIZE')
    w = read_counter(write_csv, 'WRITE_SIZE')
    cf = sum(f['k_stream_copy']) / len(f['k_stream_copy']) * 1024.0
    cw = sum(w['k_stream_copy']) / len(w['k_stream_copy']) * 1024.0
    ff, fw = copy_bytes / cf, copy_bytes / cw
    rows = []
    for k in sorted(set(f) | set(w)):
        fv, wv = f.get(k, []), w.get(k, [])
        if not fv or not wv:
            continue
        fk = sum(fv) / len(fv)
        wk = sum(wv) / len(wv)
        rows.append(dict(kernel=k, kind=kind_of(k), dispatches=len(fv), fetch_kib=fk, write_kib=wk,
                         fetch_bytes=fk * 1024.0 * ff, write_bytes=wk * 1024.0 * fw,
                         traffic_bytes=fk * 1024.0 * ff + wk * 1024.0 * fw))
    return rows, ff, fw, (cf, cw)


def traffic_by_kind(rows):
    """Mean calibrated HBM bytes per launch for each kernel kind (the stage
    variants of one kind are averaged with their dispatch counts)."""
    acc = collections.defaultdict(lambda: [0.0, 0])
    for r in rows:
        if r['kind']:
            acc[r['kind']][0] += r['traffic_bytes'] * r['dispatches']
            acc[r['kind']][1] += r['dispatches']
    return {k: v[0] / v[1] for k, v in acc.items() if v[1]}


def parse_summary(path):
    """Read back the traffic table a summary file holds: {kind: bytes per launch}."""
    out = {}
    in_table = False
    for line in open(path):
        if line.startswith('| kind |'):
            in_table = True
            continue
        if in_table:
            if not line.startswith('|'):
                break
            cells = [c.strip() for c in line.strip().strip('|').split('|')]
            if cells[0].startswith('-'):
                continue
            out[cells[0]] = float(cells[1])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--stats', required=True)
    ap.add_argument('--fetch', required=True)
    ap.add_argument('--write', required=True)
    ap.add_argument('--copy-bytes', type=float, default=float(1 << 30))
    ap.add_argument('--title', default='rocprofv3 summary')
    ap.add_argument('--command', action='append', default=[])
    a = ap.parse_args()
    stats = read_stats(a.stats)
    rows, ff, fw, (cf, cw) = traffic_table(a.fetch, a.write, a.copy_bytes)
    print(f'# {a.title}\n')
    for c in a.command:
        print(f'    {c}')
    print('\n## Kernel durations (--kernel-trace --stats)\n')
    print('| kernel | calls | average us |')
    print('|---|---|---|')
    for k, (n, ns) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        print(f'| {k} | {n} | {ns / 1e3:.2f} |')
    print('\n## HBM bytes per dispatch (--pmc FETCH_SIZE, then --pmc WRITE_SIZE)\n')
    print(f'Calibration: k_stream_copy of {a.copy_bytes:.0f} bytes reads FETCH_SIZE {cf:.0f} B and '
          f'WRITE_SIZE {cw:.0f} B per dispatch; fetch factor {ff:.4f}, write factor {fw:.4f}.\n')
    print('| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read B | HBM write B | traffic B |')
    print('|---|---|---|---|---|---|---|')
    for r in rows:
        print(f"| {r['kernel']} | {r['dispatches']} | {r['fetch_kib']:.1f} | {r['write_kib']:.1f} | "
              f"{r['fetch_bytes']:.0f} | {r['write_bytes']:.0f} | {r['traffic_bytes']:.0f} |")
    print('\n## Traffic per launch by kernel kind\n')
    print('| kind | traffic B |')
    print('|---|---|')
    for k, v in sorted(traffic_by_kind(rows).items()):
        print(f'| {k} | {v:.0f} |')


if __name__ == '__main__':
    main()
