"""Writes tests/golden/omega/phi_ac_3_energy_head.txt: the first lines of the
reference's shipped energy trace visualisation/phi_ac_3_energy.txt (data, read
as text): "iter engE" of check_soln_new for the omega run of input64.ini
(2 x 2 x 2 boxes of 68^3, ETD2, elastic, anisotropic, the sphere of
init_field_sphere_center in AC field 3).

Run once in a container that has the reference mounted:
    python tests/golden/make_omega_sphere_energy.py /root/reference
"""
import os
import sys

LINES = 11


def main(root):
    src = os.path.join(root, "visualisation", "phi_ac_3_energy.txt")
    with open(src) as f:
        head = [next(f).strip() for _ in range(LINES)]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "omega", "phi_ac_3_energy_head.txt")
    with open(dst, "w") as f:
        f.write("\n".join(head) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
