"""Summaries of rocprofv3 results databases (SQLite, rocprofv3 >= ROCm 7) for the MHD hot path.

    python -m astaroth_amd.profiling stats  <results.db>            # kernel time summary
    python -m astaroth_amd.profiling traffic <fetch.db> <write.db> <copy_bytes> <kernel-substring>

`traffic` reads FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the matching kernel, scales each
by what the same counter reads for the device-to-device copy of `copy_bytes` bytes recorded in
the same run (a kernel whose true traffic is copy_bytes read + copy_bytes written), and prints a
JSON record with HBM bytes per launch.
"""
import json
import sqlite3
import sys


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                       "max(duration) from kernels group by name order by sum(duration) desc")
    out = []
    for name, n, tot, avg, mn, mx in rows:
        out.append(dict(name=name, calls=n, total_us=tot / 1e3, avg_us=avg / 1e3,
                        min_us=mn / 1e3, max_us=mx / 1e3))
    return out


def counter_means(db, counter):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(counter_value) from pmc_events "
                       "where counter_name = ? group by name", (counter,))
    return {name: (n, val) for name, n, val in rows}


def traffic(fetch_db, write_db, copy_bytes, kernel_sub):
    f = counter_means(fetch_db, "FETCH_SIZE")
    w = counter_means(write_db, "WRITE_SIZE")

    def pick(d, sub):
        keys = [k for k in d if sub in k]
        if not keys:
            raise KeyError(f"no kernel matching {sub!r}: {sorted(d)}")
        return max(keys, key=lambda k: d[k][0])

    copy_f = pick(f, "copyBuffer")
    copy_w = pick(w, "copyBuffer")
    kf = pick(f, kernel_sub)
    kw = pick(w, kernel_sub)
    scale_f = copy_bytes / (f[copy_f][1] * 1024.0)
    scale_w = copy_bytes / (w[copy_w][1] * 1024.0)
    fetch_b = f[kf][1] * 1024.0 * scale_f
    write_b = w[kw][1] * 1024.0 * scale_w
    return dict(kernel=kf, dispatches=f[kf][0], fetch_kib_raw=f[kf][1], write_kib_raw=w[kw][1],
                copy_bytes=copy_bytes, copy_fetch_kib=f[copy_f][1], copy_write_kib=w[copy_w][1],
                fetch_scale=scale_f, write_scale=scale_w, fetch_bytes_per_launch=fetch_b,
                write_bytes_per_launch=write_b, bytes_per_launch=fetch_b + write_b)


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        for r in kernel_stats(argv[1]):
            print(f"| {r['name'][:90]} | {r['calls']} | {r['total_us']:.1f} | {r['avg_us']:.1f} | "
                  f"{r['min_us']:.1f} | {r['max_us']:.1f} |")
        return 0
    if len(argv) >= 5 and argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], int(argv[3]), argv[4]), indent=1))
        return 0
    print(__doc__)
    return 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
