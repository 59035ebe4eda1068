"""Debug verbosity thresholds (paces/config/dbg_verbosity_levels.py:19-48).

A message is printed when the Hamiltonian's ``debug_verb`` exceeds its level.
"""

INIT_VERBOSITY_LEVEL = 0
BASE_STEP_LEVEL = 1
HILBERT_SPACE_LEVEL = 2
OBSERVABLES_LEVEL = 3
HILBERT_SPACE_DETAILED_LEVEL = 4
MEM_INFO_LEVEL = 5
MOVE_DATA_TIMING_LEVEL = 5
EXPM_EVO_LEVEL = 6
SEARCHSORTED_TIMING_LEVEL = 7
