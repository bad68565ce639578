"""Writes spgrid_kat.json: SPGrid address known answers for the two GridState sizes.

The masks are derived here, independently of the oracle, from the enum formulas of
SPGrid_Mask_base / SPGrid_Mask (Lib/SPGrid/Core/SPGrid_Mask.h:22-129) with
log2_page = 12 and log2_field = log2_struct (the whole GridState, MpmGrid.h:16-36);
offsets are a parallel bit deposit of each coordinate into its mask
(Bit_Spread_u64, SPGrid_Utilities.h:81-131).  The (33,45,17) rows are the worked
examples quoted for this reference (page 3154 / 6308, element 21).
"""
import json
import os

M64 = (1 << 64) - 1


def masks(log2_struct, page_bits=12):
    data_bits = log2_struct
    block_bits = page_bits - data_bits
    zb = block_bits // 3 + (block_bits % 3 > 0)
    yb = block_bits // 3 + (block_bits % 3 > 1)
    xb = block_bits // 3
    sh = 3 - block_bits % 3 + (page_bits - 12)
    hi = (M64 << page_bits) & M64
    page_z = (0x9249249249249249 << sh) & M64 & hi
    page_y = (0x2492492492492492 << sh) & M64 & hi
    page_x = (0x4924924924924924 << sh) & M64 & hi
    ez = ((1 << zb) - 1) << data_bits
    ey = ((1 << yb) - 1) << (data_bits + zb)
    ex = ((1 << xb) - 1) << (data_bits + zb + yb)
    return page_x | ex, page_y | ey, page_z | ez


def pdep(v, mask):
    out, bit = 0, 0
    for b in range(64):
        if mask >> b & 1:
            if v >> bit & 1:
                out |= 1 << b
            bit += 1
    return out


def main():
    coords = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (33, 45, 17), (4, 4, 4), (2, 4, 4), (1023, 511, 7),
              (100, 200, 300), (4095, 4095, 4095)]
    out = {"layouts": []}
    for log2 in (6, 7):
        mx, my, mz = masks(log2)
        offs = [[list(c), hex(pdep(c[0], mx) | pdep(c[1], my) | pdep(c[2], mz))] for c in coords]
        out["layouts"].append({"log2_struct": log2, "masks": [hex(mx), hex(my), hex(mz)], "offsets": offs})
    # worked examples for this reference's layouts
    l32, l64 = out["layouts"]
    assert l32["masks"] == ["0x4924924924924c00", "0x2492492492492300", "0x92492492492490c0"]
    assert l64["masks"] == ["0x9249249249249800", "0x4924924924924600", "0x2492492492492180"]
    kat = {tuple(c): o for c, o in l32["offsets"]}
    assert kat[(1, 0, 0)] == "0x400" and kat[(0, 1, 0)] == "0x100" and kat[(0, 0, 1)] == "0x40"
    assert kat[(33, 45, 17)] == "0xc52540" and int(kat[(33, 45, 17)], 16) >> 12 == 3154
    kat = {tuple(c): o for c, o in l64["offsets"]}
    assert kat[(1, 0, 0)] == "0x800" and kat[(0, 1, 0)] == "0x200" and kat[(0, 0, 1)] == "0x80"
    assert kat[(33, 45, 17)] == "0x18a4a80" and int(kat[(33, 45, 17)], 16) >> 12 == 6308
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spgrid_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
