"""Summaries of rocprofv3 output for the committed profiles: kernel-time
tables and per-launch HBM traffic from the FETCH_SIZE / WRITE_SIZE counters,
calibrated against the library's own streaming copy kernel (k_copy16, driven
by sr_memcpy_d2d) of known size.

The counters are read from rocprofv3's ``--output-format csv`` files; this
module never touches a GPU."""
import collections
import csv
import json
import re

# kernel name (without namespace and template arguments) -> bench kernel class
_CLASSES = (
    ("k_rot2_fwd", "rot_fwd"), ("k_rot1_fwd", "rot_fwd"),
    ("k_rot2_adj", "rot_adj"), ("k_rot1_adj", "rot_adj"),
    ("k_psf_fwd", "psf_fwd"), ("k_psf_adj", "psf_adj"),
    ("k_view_reduce", "view_reduce"), ("k_img_reduce", "img_reduce"),
    ("k_atten_scan", "atten"), ("k_to_zfast", "rot_fwd"), ("k_copy16", "copy"),
)


def short_name(name):
    name = name.replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name).replace("void ", "").strip()


def kernel_class(name):
    n = short_name(name).split("::")[-1]
    for prefix, cls in _CLASSES:
        if n.startswith(prefix):
            return cls
    return None


def stats_markdown(kernel_stats_csv, command, title):
    """Markdown table of a ``--kernel-trace --stats`` kernel_stats.csv."""
    rows = list(csv.DictReader(open(kernel_stats_csv)))
    out = [f"# {title}", "", "Command (from the repository root on an MI355X):", "", "```", command, "```", "",
           "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{short_name(r['Name'])}` | {int(r['Calls'])} | {int(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {int(r['MinNs']) / 1e3:.1f} | {int(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.1f} |")
    return "\n".join(out) + "\n"


def launch_shapes_markdown(kernel_trace_csv, min_calls=8):
    """Per-(kernel, grid) duration table from a ``--kernel-trace`` CSV: the
    launch shapes of one kernel (e.g. 16-view subset launches and 64-view
    launches) averaged separately."""
    groups = collections.defaultdict(list)
    for r in csv.DictReader(open(kernel_trace_csv)):
        grid = (int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]), int(r["Grid_Size_Z"]))
        groups[(short_name(r["Kernel_Name"]), grid)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = sorted(((k, v) for k, v in groups.items() if len(v) >= min_calls), key=lambda kv: -sum(kv[1]))
    out = ["| kernel | grid (threads) | calls | avg us | median us |", "|---|---|---|---|---|"]
    for (name, grid), v in rows:
        out.append(f"| `{name}` | {grid[0]}x{grid[1]}x{grid[2]} | {len(v)} | {sum(v) / len(v) / 1e3:.1f} | "
                   f"{_median(v) / 1e3:.1f} |")
    return "\n".join(out) + "\n"


def _per_dispatch(counter_csv, counter):
    """{dispatch id: (kernel name, grid size, value)} for one counter (KiB)."""
    d = {}
    for r in csv.DictReader(open(counter_csv)):
        if r["Counter_Name"] != counter:
            continue
        k = r["Dispatch_Id"]
        name, grid, val = d.get(k, (r["Kernel_Name"], int(r["Grid_Size"]), 0.0))
        d[k] = (name, grid, val + float(r["Counter_Value"]))
    return d


def _median(v):
    v = sorted(v)
    n = len(v)
    return v[n // 2] if n % 2 else 0.5 * (v[n // 2 - 1] + v[n // 2])


def traffic(fetch_csv, write_csv, copy_bytes):
    """Per-launch HBM bytes of each kernel class.

    Each counter is scaled by (copy_bytes / what it reads for the copy
    kernel); for every class the median over its most frequent launch shape
    (kernel and grid size) is reported, which in the OS-EM benchmark is the
    one-subset launch of the kernel the timed reconstructions ran (autotuner
    candidates of the same class are other kernels or other shapes; a few
    other launches can share the shape: the median keeps them out)."""
    f = _per_dispatch(fetch_csv, "FETCH_SIZE")
    w = _per_dispatch(write_csv, "WRITE_SIZE")

    def copy_kib(d):
        v = [val for name, _, val in d.values() if kernel_class(name) == "copy"]
        if not v:
            raise ValueError("no copy-kernel dispatch in the counter file")
        return sum(v) / len(v)

    kf = copy_bytes / 1024.0 / copy_kib(f)
    kw = copy_bytes / 1024.0 / copy_kib(w)

    def by_class(d):
        out = collections.defaultdict(lambda: collections.defaultdict(list))
        for name, grid, val in d.values():
            c = kernel_class(name)
            if c and c != "copy":
                out[c][(short_name(name), grid)].append(val)
        return out

    fc, wc = by_class(f), by_class(w)
    res = {}
    for c in fc:
        key = max(fc[c], key=lambda kg: len(fc[c][kg]))
        fv, wv = fc[c][key], wc.get(c, {}).get(key, [0.0])
        fb = 1024.0 * kf * _median(fv)
        wb = 1024.0 * kw * _median(wv)
        res[c] = {"kernel": key[0], "grid_size": key[1], "dispatches": len(fv), "fetch_bytes": round(fb),
                  "write_bytes": round(wb), "traffic_bytes": round(fb + wb)}
    return {"calibration": {"copy_bytes": copy_bytes, "fetch_scale": kf, "write_scale": kw}, "kernels": res}


def write_traffic(fetch_csv, write_csv, copy_bytes, out_json):
    t = traffic(fetch_csv, write_csv, copy_bytes)
    with open(out_json, "w") as fh:
        json.dump(t, fh, indent=1, sort_keys=True)
    return t
