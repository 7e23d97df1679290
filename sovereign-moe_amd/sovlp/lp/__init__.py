"""LP engines: single-LP PDLP, batched PDLP, mixed precision with refinement."""
