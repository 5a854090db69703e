"""Extract the built-in element table of tpt-core as data.

Reads the reference's element definitions as *text* (nothing is compiled or
imported) and writes ``elements.json``: one record per element slot 0..255
with the physics fields that ``GpuElementProps`` carries
(``crates/tpt-core/src/elements/mod.rs:482-521``).

Sources it follows:

* ``crates/tpt-core/src/constants.rs`` -- ``PT_*``, ``TYPE_*``, ``PROP_*``,
  ``R_TEMP``, ``IPL``/``IPH``/``ITL``/``ITH``, ``NT``/``ST``;
* ``crates/tpt-core/src/elements/mod.rs:186-236`` -- ``ElementDef::default``
  (the values every ``elem!`` record and every padding slot start from);
* ``crates/tpt-core/src/elements/mod.rs:942-973`` -- the ``elem!`` record
  grammar (``id, ident, name, cat, [photon_reflect], colour, props, weight,
  desc, field: value...``);
* ``crates/tpt-core/src/elements/data/*.rs`` -- the 240 records;
* ``crates/tpt-core/src/elements/mod.rs:307-340`` -- ``carries_ctype``.

The rendering-only words (``pixel_mode``, ``fire_colour``) are copied as
written in the record; the reference's ``apply_default_graphics`` rewrite of
them is not reproduced because no simulation kernel reads them.

Run:  python gen_elements.py /root/reference  (writes elements.json beside it)
"""

import json
import os
import re
import sys

FIELDS_F32 = {
    "advection", "air_drag", "air_loss", "loss", "collision", "gravity",
    "diffusion", "hot_air", "default_temp", "low_pressure", "high_pressure",
    "low_temp", "high_temp",
}


def _strip_comments(src):
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.sub(r"//[^\n]*", "", src)


def parse_constants(path):
    """Evaluate the numeric ``pub const`` items of constants.rs in order."""
    env = {}
    src = _strip_comments(open(path).read())
    items = []
    for m in re.finditer(r"pub const (\w+)\s*:\s*([\w\[\]; ]+?)\s*=\s*(.+?);", src, flags=re.S):
        name, ty, expr = m.group(1), m.group(2).strip(), m.group(3)
        if ty in ("u32", "i32", "f32", "usize", "u8", "u16", "i64", "u64"):
            items.append((name, expr))
    # Items may refer to later ones (PT_NUM uses BUILTIN_ELEMENT_COUNT).
    for _ in range(4):
        for name, expr in items:
            if name in env:
                continue
            try:
                env[name] = _eval(expr, env)
            except Exception:
                pass
    return env


def _eval(expr, env):
    e = " ".join(expr.split())
    e = re.sub(r"\bas\s+(u32|i32|f32|usize|u8|u16|i64|u64|f64)\b", "", e)
    e = re.sub(r"\b(0x[0-9a-fA-F_]+|[0-9][0-9_]*)", lambda m: m.group(1).replace("_", ""), e)
    e = re.sub(r"(\d)(f32|f64|u32|i32|usize)\b", r"\1", e)
    e = e.replace("!", "~")
    return eval(e, {"__builtins__": {}}, dict(env))


def split_top_level(body):
    """Split an ``elem!`` body into ``key: value`` pairs at depth-0 commas."""
    out, depth, cur, in_str = [], 0, "", False
    i = 0
    while i < len(body):
        c = body[i]
        if in_str:
            cur += c
            if c == "\\":
                cur += body[i + 1]
                i += 1
            elif c == '"':
                in_str = False
        elif c == '"':
            in_str = True
            cur += c
        elif c in "([{":
            depth += 1
            cur += c
        elif c in ")]}":
            depth -= 1
            cur += c
        elif c == "," and depth == 0:
            if cur.strip():
                out.append(cur.strip())
            cur = ""
        else:
            cur += c
        i += 1
    if cur.strip():
        out.append(cur.strip())
    pairs = []
    for item in out:
        k, v = item.split(":", 1)
        pairs.append((k.strip(), v.strip()))
    return pairs


def find_records(src):
    """Yield the body text of every ``elem!( ... )`` invocation."""
    pos = 0
    while True:
        i = src.find("elem!(", pos)
        if i < 0:
            return
        j = i + len("elem!(")
        depth, in_str = 1, False
        k = j
        while depth:
            c = src[k]
            if in_str:
                if c == "\\":
                    k += 1
                elif c == '"':
                    in_str = False
            elif c == '"':
                in_str = True
            elif c == "(":
                depth += 1
            elif c == ")":
                depth -= 1
            k += 1
        yield src[j:k - 1]
        pos = k


def default_record(env, slot):
    """``ElementDef::default`` (elements/mod.rs:186-236) for slot ``slot``."""
    return {
        "id": slot, "identifier": "DEFAULT_INVALID", "name": "", "enabled": False,
        "colour": 0xFF00FF, "properties": env["TYPE_SOLID"], "photon_reflect": 0x3FFFFFFF,
        "weight": 50, "advection": 0.0, "air_drag": 0.0, "air_loss": 1.0, "loss": 1.0,
        "collision": 0.0, "gravity": 0.0, "diffusion": 0.0, "hot_air": 0.0, "falldown": 0,
        "flammable": 0, "explosive": 0, "meltable": 0, "hardness": 30, "heat_conduct": 128,
        "default_temp": env["R_TEMP"] + 273.15, "default_life": 0, "default_ctype": 0,
        "default_tmp": 0, "default_tmp2": 0, "default_tmp3": 0, "default_tmp4": 0,
        "pixel_mode": 0, "fire_colour": 0,
        "low_pressure": env["IPL"], "low_pressure_transition": env["NT"],
        "high_pressure": env["IPH"], "high_pressure_transition": env["NT"],
        "low_temp": env["ITL"], "low_temp_transition": env["NT"],
        "high_temp": env["ITH"], "high_temp_transition": env["NT"],
    }


def main(ref_root):
    core = os.path.join(ref_root, "crates", "tpt-core", "src")
    env = parse_constants(os.path.join(core, "constants.rs"))
    elem_mod = _strip_comments(open(os.path.join(core, "elements", "mod.rs")).read())
    for m in re.finditer(r"const (PMODE\w*|NO_DECO|DECO_FIRE|FIRE_\w+)\s*:\s*u32\s*=\s*(.+?);", elem_mod):
        env[m.group(1)] = _eval(m.group(2), env)
    pt_num = env["PT_NUM"]
    table = [default_record(env, s) for s in range(pt_num)]
    # Slot 0 is written as a struct literal, not an elem! record
    # (elements/data/special.rs:13-30).
    table[0].update({"identifier": "DEFAULT_PT_NONE", "name": "NONE", "colour": 0,
                     "properties": 0, "weight": 0, "heat_conduct": 0, "default_temp": 295.15})

    data_dir = os.path.join(core, "elements", "data")
    seen = 0
    for fname in sorted(os.listdir(data_dir)):
        if not fname.endswith(".rs"):
            continue
        src = _strip_comments(open(os.path.join(data_dir, fname)).read())
        for body in find_records(src):
            if "id:" not in body:
                continue
            pairs = split_top_level(body)
            rec = None
            for k, v in pairs:
                if k == "id":
                    slot = int(_eval(v, env))
                    rec = default_record(env, slot)
                    rec["enabled"] = True
                elif k == "ident":
                    rec["identifier"] = json.loads(v)
                elif k == "name":
                    rec["name"] = json.loads(v)
                elif k in ("desc", "cat", "ctype_draw"):
                    continue
                elif k == "props":
                    rec["properties"] = int(_eval(v, env))
                else:
                    val = _eval(v, env)
                    rec[k] = float(val) if k in FIELDS_F32 else int(val)
            table[rec["id"]] = rec
            seen += 1

    m = re.search(r"pub fn carries_ctype\(id: u32\) -> bool \{\s*matches!\(\s*id,(.*?)\)", elem_mod, flags=re.S)
    carries = {env[n.strip()] for n in m.group(1).split("|") if n.strip()}
    for rec in table:
        rec["carries_ctype"] = 1 if rec["id"] in carries else 0

    out = {
        "source": "crates/tpt-core/src/elements (text-extracted)",
        "pt_num": pt_num,
        "builtin_count": env["BUILTIN_ELEMENT_COUNT"],
        "records_parsed": seen,
        "constants": {k: v for k, v in env.items() if re.match(r"^(PT_|TYPE_|PROP_|FLAG_|WL_)", k)},
        "elements": table,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elements.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print(f"wrote {dst}: {seen} records, PT_NUM={pt_num}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
