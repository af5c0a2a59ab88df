"""VALU instructions per DP cell in the inner loop of each kernel.

Compiles csrc/swg_kernels.hip for gfx950 to assembly (this build's own
code), takes each kernel's largest basic block (the unrolled 4-column x
R-row loop body, with the block that closes the loop when the compiler
splits it in two) and divides its VALU instruction count by the cells that
body evaluates.  The score pass folds its block maxima into the row maxima
once per 16-column block, i.e. after every fourth body: the instructions the
body skips with a conditional branch count a quarter.  bench.py's roofline
uses these figures as the kernels' operations per cell.

usage: python -m csharp_amd.isa_valu_count
"""
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

# kernel symbol fragment -> DP cells evaluated by one pass of the loop body
# (and, optionally, the passes per execution of its conditional tail)
KERNELS = {
    "score16_kernelILi24ELb1E": ("score16", 4 * 24 * 2, 4),    # 4 columns x 24 rows x 2 pairs per lane
    "score16_kernelILi24ELb0E": ("score16 (large alphabet)", 4 * 24 * 2, 4),
    "rerun_kernelILi16ELb0E": ("rerun_capture", 4 * 16),
    "reverse_kernelILi16E": ("reverse", 4 * 16 * 2),           # 2 pairs per lane
    "sw_kernelILi16ELb1E": ("one_pass", 4 * 16),
    "nw_kernelILi16ELb1E": ("global", 16),
}


def count(asm_text):
    out = {}
    for frag, spec in KERNELS.items():
        name, cells = spec[0], spec[1]
        every = spec[2] if len(spec) > 2 else 1
        m = re.search(r"\n(_Z\S*%s\S*):[^\n]*\n" % re.escape(frag), asm_text)
        if not m:
            continue
        start = m.end()
        end = asm_text.index(".Lfunc_end", start)
        body = asm_text[start:end]
        parts = re.split(r"\n(\.LBB[0-9_]+):", body)
        blocks = {parts[k]: parts[k + 1] for k in range(1, len(parts), 2)}

        def valu(txt):
            return sum(1 for line in txt.split("\n") if line.strip().startswith("v_"))

        def branch_target(txt):
            lines = [ln.strip() for ln in txt.split("\n") if ln.strip() and not ln.strip().startswith((";", "."))]
            m2 = re.match(r"s_branch\s+(\.LBB[0-9_]+)", lines[-1]) if lines else None
            return m2.group(1) if m2 else None

        label = max(blocks, key=lambda lb: valu(blocks[lb]))
        best = valu(blocks[label])
        if every > 1:
            # the part after the body's last conditional forward branch runs
            # once per `every` passes
            m3 = list(re.finditer(r"\n\s*s_cbranch_\w+\s+\.LBB", blocks[label]))
            if m3:
                tail = valu(blocks[label][m3[-1].end():])
                best = best - tail + tail / every
        # a loop body the compiler closed with a second block that jumps back
        nxt = branch_target(blocks[label])
        if nxt in blocks and nxt != label and re.search(r"s_c?branch\S*\s+%s\b" % re.escape(label), blocks[nxt]):
            best += valu(blocks[nxt])
        out[name] = round(best / cells, 3)
    return out


def main():
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", "-I" + os.path.join(HERE, "csrc")]
    with tempfile.TemporaryDirectory() as d:
        s = os.path.join(d, "k.s")
        subprocess.run(["/opt/rocm/bin/hipcc"] + flags + [os.path.join(HERE, "csrc", "swg_kernels.hip"), "-o", s],
                       check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        print(count(open(s).read()))


if __name__ == "__main__":
    main()
