"""Degenerate instances shared by the decoder and block-engine parity tests
(thpack "original" text: count, id seed, L W H, n types, then per type
id l lf w wf h hf qty)."""

EDGE_TEXTS = {
    # no orientation of any item fits the container: nothing is placed
    "nofit": "1\n1 7\n100 100 100\n2\n1 150 1 150 1 150 1 3\n2 120 1 20 1 20 1 2\n",
    # one item
    "single": "1\n1 8\n100 100 100\n1\n1 30 1 40 1 50 1 1\n",
    # a type with no allowed orientation among placeable ones
    "someflags0": "1\n1 9\n200 150 100\n3\n1 30 0 40 0 50 0 2\n2 30 1 40 1 50 1 3\n"
                  "3 60 0 20 1 30 0 2\n",
    # no type has an allowed orientation
    "allflags0": "1\n1 10\n100 100 100\n1\n1 10 0 10 0 10 0 5\n",
    # boxes that tile the container exactly
    "exact": "1\n1 11\n60 40 30\n1\n1 20 1 20 1 30 1 6\n",
}
