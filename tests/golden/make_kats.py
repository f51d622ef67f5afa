"""Write tests/golden/kats.json: the known-answer vectors of SReD's own JUnit
tests for the 2D path, transcribed as data (inputs, expected outputs and the
tolerance each reference test uses), with the file:line they come from.

    python tests/golden/make_kats.py

Paths are relative to the SReD repository root.  Nothing here runs or
imports the reference; the values are the literals its tests assert.
"""
import json
import os

CROSS_10x10 = [[1.0 if (x == 5 or y == 4) else 0.0 for x in range(10)] for y in range(10)]
VBAR_3x3 = [[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 1.0, 0.0]]

_Z = 0.0
KATS = {
    "local_statistics_2d": {
        "source": "src/test/java/CLUtilsTest.java:455-527",
        "image": [[float(5 * y + x + 1) for x in range(5)] for y in range(5)],
        "block": [3, 3],
        "expected_means": [[0, 0, 0, 0, 0], [0, 7, 8, 9, 0], [0, 12, 13, 14, 0], [0, 17, 18, 19, 0], [0, 0, 0, 0, 0]],
        "expected_stds": [[0, 0, 0, 0, 0]] + [[0, 3.6055512, 3.6055512, 3.6055512, 0]] * 3 + [[0, 0, 0, 0, 0]],
        "tol": 1e-7,
    },
    # CLUtilsTest builds ReferenceBlock2D directly from the raw (not
    # mean-subtracted) footprint pixels, their mean and their unbiased std
    # (CLUtilsTest.java:560-597); relevance constant 0, normalise on.
    "block_pearson_2d": {
        "source": "src/test/java/CLUtilsTest.java:531-634",
        "image": CROSS_10x10, "block": VBAR_3x3, "metric": "Pearson's R", "tol": 1e-6,
        "expected": [
            [_Z] * 10,
            [0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0],
            [0, 0.40824822, 0.40824822, 0.40824822, 0, 1.0, 0, 0.40824822, 0.40824822, 0],
            [_Z] * 10,
            [0, 0.40824822, 0.40824822, 0.40824822, 0, 1.0, 0, 0.40824822, 0.40824822, 0],
            [0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0],
            [_Z] * 10,
        ],
    },
    "block_ssim_2d": {
        "source": "src/test/java/CLUtilsTest.java:641-746",
        "image": CROSS_10x10, "block": VBAR_3x3, "metric": "SSIM", "tol": 1e-7,
        "expected": [
            [_Z] * 10,
            [0, 0, 0, 0, 0, 0.9999999, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 0.9999999, 0, 0, 0, 0],
            [0, 0.24068691, 0.24068691, 0.24068691, 0, 0.9999999, 0, 0.24068691, 0.24068691, 0],
            [_Z] * 10,
            [0, 0.24068692, 0.24068692, 0.24068692, 0, 0.9999999, 0, 0.24068692, 0.24068692, 0],
            [0, 0, 0, 0, 0, 0.9999999, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 0.9999999, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 0.9999999, 0, 0, 0, 0],
            [_Z] * 10,
        ],
    },
    "block_nrmse_2d": {
        "source": "src/test/java/CLUtilsTest.java:750-857",
        "image": CROSS_10x10, "block": VBAR_3x3, "metric": "NRMSE (inverted)", "tol": 1e-5,
        "expected": [
            [_Z] * 10,
            [0, 0.99999994, 0.99999994, 0.99999994, 0.11117814, 0, 0.11117814, 0.99999994, 0.99999994, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.11117814, 0, 0.11117814, 0.99999994, 0.99999994, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.18945144, 0, 0.18945158, 0.99999994, 0.99999994, 0],
            [0, 1.3763912e-7, 1.3763912e-7, 1.3763912e-7, 1.3763912e-7, 0.99999994, 1.3763912e-7, 1.3763912e-7,
             1.3763912e-7, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.18945144, 0, 0.18945144, 0.99999994, 0.99999994, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.11117814, 0, 0.11117814, 0.99999994, 0.99999994, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.11117814, 0, 0.11117814, 0.99999994, 0.99999994, 0],
            [0, 0.99999994, 0.99999994, 0.99999994, 0.11117814, 0, 0.11117814, 0.99999994, 0.99999994, 0],
            [_Z] * 10,
        ],
    },
    "block_absdiff_2d": {
        "source": "src/test/java/CLUtilsTest.java:861-966",
        "image": CROSS_10x10, "block": VBAR_3x3, "metric": "Abs. Diff. of StdDevs.", "tol": 1e-7,
        "expected": [
            [_Z] * 10,
            [0, 0, 0, 0, 0.8164966, 0.9999999, 0.8164966, 0, 0, 0],
            [0, 0, 0, 0, 0.8164966, 0.9999999, 0.8164966, 0, 0, 0],
            [0, 0.8164966, 0.8164966, 0.8164966, 0.99999976, 0.9999999, 0.99999976, 0.8164966, 0.8164966, 0],
            [0, 0.9999999, 0.9999999, 0.9999999, 0.9999999, 0, 0.9999999, 0.9999999, 0.9999999, 0],
            [0, 0.8164966, 0.8164966, 0.8164966, 0.99999976, 0.9999999, 0.99999976, 0.8164966, 0.8164966, 0],
            [0, 0, 0, 0, 0.8164966, 0.9999999, 0.8164966, 0, 0, 0],
            [0, 0, 0, 0, 0.8164966, 0.9999999, 0.8164966, 0, 0, 0],
            [0, 0, 0, 0, 0.8164966, 0.9999999, 0.8164966, 0, 0, 0],
            [_Z] * 10,
        ],
    },
    "reference_block_2d": {
        "source": "src/test/java/UtilsTest.java:296-336",
        "block": [[1, 2, 3, 4, 5], [2, 3, 4, 5, 6], [3, 4, 5, 6, 7], [4, 5, 6, 7, 8], [5, 6, 7, 8, 9]],
        "expected_mean": 0.5, "expected_std": 0.64549720287323, "tol": 1e-6,
        "even_block": [[1, 2, 3, 4], [2, 3, 4, 5], [3, 4, 5, 6], [4, 5, 6, 7]],
    },
    "input_image_2d_normalize": {
        "source": "src/test/java/UtilsTest.java:411-447",
        "image": [[10.0, 20.0], [30.0, 40.0]],
        "expected": [[0.0, 0.3333], [0.6667, 1.0]], "tol": 1e-4,
    },
    "mean_var_block_2d": {
        "source": "src/test/java/UtilsTest.java:556-601",
        "image_rule": "x + y on 100x100", "block": [10, 10, 20, 20],
        "expected_mean": 29.0, "expected_var": 16.5, "tol": 0.01,
    },
    "apply_mask_2d": {
        "source": "src/test/java/UtilsTest.java:611-641",
        "image": [[1, 2, 3], [4, 5, 6], [7, 8, 9]],
        "mask": [[1, 0, 1], [0, 1, 0], [1, 1, 0]],
        "expected": [[1, 0, 3], [0, 5, 0], [7, 8, 0]],
    },
    "relevance_mask_random42": {
        "source": "src/test/java/UtilsTest.java:645-690",
        "width": 50, "height": 50, "seed": 42, "block_radius": [2, 2], "relevance_constant": 2.0,
        "fill_rule": "for i: img[i] = i%50==0 ? 1 : 1 + nextFloat()*2; stds[i] = i%50==0 ? 0 : nextFloat()",
        "expected_threshold_f32": 0.4694742262363434, "expected_noise_var": 0.2347371131181717,
    },
    "normalize_image_2d_mask": {
        "source": "src/test/java/UtilsTest.java:696-731",
        "image": [[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12], [13, 14, 15, 16]],
        "mask": [[1, 0, 1, 0], [1, 1, 0, 0], [1, 0, 1, 1], [0, 0, 0, 1]],
        "border": [1, 1],
        "expected": [[1, 2, 3, 4], [5, 0, 7, 8], [9, 10, 1, 12], [13, 14, 15, 16]], "tol": 1e-4,
    },
    "normalize_image_2d_nonbinary": {
        "source": "src/test/java/UtilsTest.java:734-765",
        "image": [[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12], [13, 14, 15, 16]],
        "mask": [[1, 0.5, 1, 0], [1, 0.8, 0, 0], [1, 0, 1, 1], [0, 0, 0, 1]],
        "border": [0, 0], "expect_error": True,
    },
    "block_size_table": {
        "source": "GlobalRepetition2D_.java:102-111; circular blocks count the lattice points of x^2+y^2<=r^2 (Gauss circle numbers, OEIS A000328)",
        "sizes": {"3": 5, "5": 13, "9": 49, "11": 81, "13": 113, "15": 149, "17": 197},
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print(out)
