"""MI355X-native FluidiTy liquid-step hot path: P2G, G2P and MAC pressure projection (MGPCG).

The compute path is the HIP library fluidity_amd/libfluidity_hip.so behind the C-ABI declared in
include/fluidity_hip.h; `fluidity_amd.hip` binds it with ctypes and raises if it is missing.
`records` holds the C-ABI record layouts and `scenes` the seeded synthetic inputs.
"""
