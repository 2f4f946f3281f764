"""Copy kete's SimultaneousStates bincode fixture (src/tests/data/
simult_state_v0.2.1.bin, loaded by src/tests/test_state.py:62-65) to
tests/golden/ as data.  Run in a checkout that holds the kete sources:
python tests/golden/copy_simult_state_fixture.py [path/to/simult_state_v0.2.1.bin]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else \
    "/root/reference/src/tests/data/simult_state_v0.2.1.bin"

if __name__ == "__main__":
    shutil.copyfile(SRC, os.path.join(HERE, "simult_state_v0.2.1.bin"))
