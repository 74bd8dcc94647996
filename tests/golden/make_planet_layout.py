"""Writes planet_layout_earth_g4.json: the dataset layout of the planet file
THOR ships with its Earth Held-Suarez input (ifile/esp_initial_planet.h5:
glevel 4, vlevel 32), as read by this project's HDF5 reader. Each dataset's
integer/float type, size, "Variable" and "units" attributes and, for
scalars, its value. Run in a tree where the reference is present:

    python tests/golden/make_planet_layout.py /root/reference/ifile/esp_initial_planet.h5
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from thor_amd import io  # noqa: E402

if __name__ == "__main__":
    layout = io.h5_layout(sys.argv[1])
    with open(os.path.join(HERE, "planet_layout_earth_g4.json"), "w") as fh:
        json.dump(layout, fh, indent=1, sort_keys=True)
    print(f"{len(layout)} datasets")
