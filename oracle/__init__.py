"""The oracle: CPU restatement of tblite's GFN2 integral / H0 / gradient path.

TEST INFRASTRUCTURE ONLY -- see ``oracle.py`` for who may import it.
"""
