"""bms2_amd: MI355X-native batched subgraph isomorphism (BMS2 hot path) behind a C ABI."""
