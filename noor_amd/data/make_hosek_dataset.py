"""Writes hosek_rgb_dataset.json: the coefficient tables of the RGB Hosek-Wilkie
sky model (Hosek & Wilkie 2012, "An Analytic Model for Full Spectral Sky-Dome
Radiance", 3-clause BSD licence, copyright 2012-2013 Lukas Hosek and
Alexander Wilkie), read as text from the reference's vendored copy
(noor/external/ArHosekSkyModelData_RGB.h).

Only the numbers are kept, as the decimal strings of the table (so that a
C compiler turns each into the same binary32 value as the original source):
per channel, 2 albedos x 10 turbidities x 6 control points x 9 parameters
configuration coefficients, and 2 x 10 x 6 radiance coefficients.

Run once in a checkout that has the reference at /root/reference:
    python noor_amd/data/make_hosek_dataset.py
"""
import json
import os
import re

SRC = "/root/reference/noor/external/ArHosekSkyModelData_RGB.h"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "hosek_rgb_dataset.json")

LICENCE = (
    "Coefficient data of the RGB Hosek-Wilkie sky model, version 1.4a. Copyright (c) 2012 - 2013, "
    "Lukas Hosek and Alexander Wilkie. All rights reserved. Redistribution and use in source and binary "
    "forms, with or without modification, are permitted provided that the following conditions are met: "
    "Redistributions of source code must retain the above copyright notice, this list of conditions and "
    "the following disclaimer. Redistributions in binary form must reproduce the above copyright notice, "
    "this list of conditions and the following disclaimer in the documentation and/or other materials "
    "provided with the distribution. None of the names of the contributors may be used to endorse or "
    "promote products derived from this software without specific prior written permission. THIS "
    "SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS \"AS IS\" AND ANY EXPRESS OR IMPLIED "
    "WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR "
    "A PARTICULAR PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDERS BE LIABLE FOR ANY "
    "DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT "
    "LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS "
    "INTERRUPTION) HOWEVER CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, "
    "OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF THIS SOFTWARE, EVEN "
    "IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.")


def table(text, name):
    m = re.search(r"float\s+" + name + r"\[\]\s*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [t for t in (x.strip() for x in body.split(",")) if t]


def main():
    text = open(SRC).read()
    cfg = [table(text, f"datasetRGB{c}") for c in (1, 2, 3)]
    rad = [table(text, f"datasetRGBRad{c}") for c in (1, 2, 3)]
    assert all(len(t) == 2 * 10 * 6 * 9 for t in cfg) and all(len(t) == 2 * 10 * 6 for t in rad)
    for t in cfg + rad:
        for x in t:
            float(x)
    doc = {
        "description": "RGB Hosek-Wilkie sky model coefficients: configs[channel] holds 2 albedos x 10 "
                       "turbidities x 6 quintic Bezier control points x 9 parameters; radiances[channel] "
                       "2 x 10 x 6. Decimal strings as published.",
        "licence": LICENCE,
        "configs": cfg,
        "radiances": rad,
    }
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=0)
        fh.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
