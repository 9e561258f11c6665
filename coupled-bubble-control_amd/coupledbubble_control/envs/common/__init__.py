"""Action / observation schemas and spaces of the vector environment
(reference: coupledbubble_control/envs/common/{specs,spaces}.py)."""
